"""Synthetic workloads of the benchmark configurations, as pmcx dicts.

cube60: 60x60x60 uint8, all label 1, 1 mm voxels, medium 1 (0.005/mm,
  1.0/mm, g 0.01, n 1.37), pencil at (29,29,0) along +z, one 0-5 ns gate,
  seed 1648335518, reflection on (the reference's cube60b bench,
  src/mcx_bench.c:105-150, without detectors).
sphere100: 100^3 uint8, background label 1 (0.005, 1.0, 0.9, 1.37) with a
  centred sphere of radius 15 voxels as label 2 (0.05, 5.0, 0.9, 1.37),
  outside n=1 with Fresnel boundaries, pencil at (50,50,0) along +z,
  50 gates of 0.1 ns, 4 disk detectors of radius 2 on z=0 at 10/20/30/40 mm
  from the source saving detector id and partial paths, seed 12345.
"""
import numpy as np


def cube60(nphoton=1_000_000, reflect=1, seed=1648335518, **kw):
    cfg = {
        "nphoton": int(nphoton),
        "vol": np.ones((60, 60, 60), dtype=np.uint8),
        "prop": np.array([[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37]], dtype=np.float32),
        "srcpos": [29, 29, 0],
        "srcdir": [0, 0, 1],
        "issrcfrom0": 1,
        "tstart": 0.0,
        "tend": 5e-9,
        "tstep": 5e-9,
        "seed": seed,
        "isreflect": reflect,
        "issavedet": 0,
    }
    cfg.update(kw)
    return cfg


def sphere_volume(n=100, radius=15.0):
    """Centred sphere by the Sphere shape rule (voxel centres i+0.5 within R,
    src/mcx_shapes.c:237-288); label 2 inside, 1 outside, x fastest."""
    c = np.arange(n, dtype=np.float32) + np.float32(0.5)
    o = np.float32(n / 2.0)
    dx = (c - o)[:, None, None]
    dy = (c - o)[None, :, None]
    dz = (c - o)[None, None, :]
    r2 = dx * dx + dy * dy + dz * dz
    vol = np.ones((n, n, n), dtype=np.uint8)
    vol[r2 <= np.float32(radius * radius)] = 2
    return vol


def sphere100(nphoton=100_000_000, seed=12345, **kw):
    cfg = {
        "nphoton": int(nphoton),
        "vol": sphere_volume(100, 15.0),
        "prop": np.array([[0, 0, 1, 1], [0.005, 1.0, 0.9, 1.37], [0.05, 5.0, 0.9, 1.37]], dtype=np.float32),
        "srcpos": [50, 50, 0],
        "srcdir": [0, 0, 1],
        "issrcfrom0": 1,
        "tstart": 0.0,
        "tend": 5e-9,
        "tstep": 1e-10,
        "seed": seed,
        "isreflect": 1,
        "detpos": np.array([[60, 50, 0, 2], [70, 50, 0, 2], [80, 50, 0, 2], [90, 50, 0, 2]], dtype=np.float32),
        "savedetflag": "DP",
        "issavedet": 1,
    }
    cfg.update(kw)
    return cfg


# Published absorbed fractions of the reference's three speed-contest
# problems at 1e8 photons (speedcontest/mcxcontest:103-105;
# mcxlab/examples/mcx_gpu_benchmarks.m:65, :87, :112 with seed 29012392).
# cubesph60b is 0.2701 in the mcxlab script and 0.2702 in the contest script.
CONTEST_ABSORBED = {"cube60": 0.1769, "cubesph60b": 0.2701, "cube60planar": 0.2551}


def contest(name, nphoton=100_000_000, seed=29012392, **kw):
    """The reference's speed-contest problems as the mcxlab benchmark script
    defines them (mcxlab/examples/mcx_gpu_benchmarks.m:46-112, same as the
    CLI bench JSON of src/mcx_bench.c apart from the seed): a 60^3 cube of
    label 1 (0.005/mm, 1.0/mm, g 0.01, n 1.37), a third medium (0.002/mm,
    5.0/mm, g 0.9, n 1.0), four r=1 detectors on z=0, one 0-5 ns gate.
      cube60:       pencil at (29,29,0) along +z, no reflection;
      cubesph60b:   plus a sphere of label 2 centred at 30 with R=15 (index
                    1.0 inside 1.37: an internal Fresnel interface and
                    g=0.9), reflection on;
      cube60planar: no sphere, a 40x40 planar source at (10,10,-10) in the
                    void above the cube, reflection on.
    """
    cfg = {
        "nphoton": int(nphoton),
        "vol": np.ones((60, 60, 60), dtype=np.uint8),
        "prop": np.array([[0, 0, 1, 1], [0.005, 1.0, 0.01, 1.37], [0.002, 5.0, 0.9, 1.0]], dtype=np.float32),
        "srcpos": [29, 29, 0],
        "srcdir": [0, 0, 1],
        "issrcfrom0": 1,
        "tstart": 0.0,
        "tend": 5e-9,
        "tstep": 5e-9,
        "seed": seed,
        "detpos": np.array([[29, 19, 0, 1], [29, 39, 0, 1], [19, 29, 0, 1], [39, 29, 0, 1]], dtype=np.float32),
        "issavedet": 1,
        "isreflect": 0,
    }
    if name == "cubesph60b":
        cfg["vol"] = sphere_volume(60, 15.0)
        cfg["isreflect"] = 1
    elif name == "cube60planar":
        cfg.update(srctype="planar", srcparam1=[40, 0, 0, 0], srcparam2=[0, 40, 0, 0], srcpos=[10, 10, -10],
                   isreflect=1)
    elif name != "cube60":
        raise ValueError("unknown contest problem %r" % name)
    cfg.update(kw)
    return cfg


def batch_instance(master_seed, index, edge_range=(40, 160), photon_range=(1e5, 1e7), ngates=10, tstep=5e-10):
    """One instance of the heterogeneous batch workload.

    A counter-based generator (Philox keyed by the master seed and the
    instance index) draws, in this order: the cube edge (uniform integer in
    edge_range), 2-6 planar layers along z (cut planes uniform, distinct),
    0-3 ellipsoidal inclusions (semi-axes 3-20 voxels, centres uniform in
    the cube), per-label optical properties (mua U(0.001, 0.1)/mm, mus
    U(0.5, 20)/mm, g U(0.7, 0.95), n U(1.33, 1.45)), the photon count
    (log-uniform in photon_range) and the RNG seed.  Pencil source at the
    centre of the z=0 face along +z; 10 gates of 0.5 ns; reflection on.
    """
    g = np.random.Generator(np.random.Philox(key=[int(master_seed), int(index)]))
    edge = int(g.integers(edge_range[0], edge_range[1] + 1))
    nlayer = int(g.integers(2, 7))
    cuts = np.sort(g.choice(np.arange(1, edge), size=nlayer - 1, replace=False))
    z = np.arange(edge)
    zlab = (1 + np.searchsorted(cuts, z, side="right")).astype(np.uint8)
    vol = np.broadcast_to(zlab[None, None, :], (edge, edge, edge)).copy()
    nell = int(g.integers(0, 4))
    c = np.arange(edge, dtype=np.float32) + np.float32(0.5)
    for k in range(nell):
        ctr = g.uniform(0, edge, size=3).astype(np.float32)
        ax = g.uniform(3, 20, size=3).astype(np.float32)
        r = (((c - ctr[0]) / ax[0])[:, None, None] ** 2 + ((c - ctr[1]) / ax[1])[None, :, None] ** 2
             + ((c - ctr[2]) / ax[2])[None, None, :] ** 2)
        vol[r <= 1.0] = nlayer + 1 + k
    nlab = nlayer + nell
    prop = np.zeros((nlab + 1, 4), dtype=np.float32)
    prop[0] = [0, 0, 1, 1]
    prop[1:, 0] = g.uniform(0.001, 0.1, size=nlab)
    prop[1:, 1] = g.uniform(0.5, 20.0, size=nlab)
    prop[1:, 2] = g.uniform(0.7, 0.95, size=nlab)
    prop[1:, 3] = g.uniform(1.33, 1.45, size=nlab)
    lo, hi = np.log(photon_range[0]), np.log(photon_range[1])
    nphoton = int(round(float(np.exp(g.uniform(lo, hi)))))
    seed = int(g.integers(1, 2**31 - 1))
    return {
        "nphoton": nphoton,
        "vol": vol,
        "prop": prop,
        "srcpos": [edge / 2.0, edge / 2.0, 0],
        "srcdir": [0, 0, 1],
        "issrcfrom0": 1,
        "tstart": 0.0,
        "tend": tstep * ngates,
        "tstep": tstep,
        "seed": seed,
        "isreflect": 1,
        "issavedet": 0,
    }


def batch256(master_seed=7, n=256, **kw):
    """The 256-instance heterogeneous batch (edges 40-160, 1e5-1e7 photons)."""
    return [batch_instance(master_seed, i, **kw) for i in range(n)]


def _perturbed_radius(g, theta, phi, amp, lmax=4):
    """1 + a smooth random radial perturbation of relative amplitude ~amp:
    a sum of low-order products cos(l theta + a) cos(m phi + b)."""
    r = np.ones_like(theta)
    for l in range(1, lmax + 1):
        for m in range(0, lmax + 1):
            c = g.normal(0, amp / (l + m + 1))
            r = r + c * np.cos(l * theta + g.uniform(0, 2 * np.pi)) * np.cos(m * phi + g.uniform(0, 2 * np.pi))
    return r


def head256(nphoton=1_000_000_000, seed=99, n=256, unitinmm=0.5, **kw):
    """Large-scale five-label head-like volume: nested randomly perturbed
    ellipsoidal shells on a 256^3 grid of 0.5 mm voxels.  Shell thicknesses
    (mm): outer 3-5, next 5-8, next 1-3; the core is split into two random
    regions by a smooth random surface.  Media (per mm): 1 outer (0.019, 7.8,
    0.89, 1.37), 2 next (0.019, 7.8, 0.89, 1.37), 3 thin layer (0.004,
    0.009, 0.89, 1.37), 4 core A (0.02, 9.0, 0.89, 1.37), 5 core B (0.08,
    40.9, 0.89, 1.37).  Pencil source entering the outer surface at the top
    of the head along -z; 20 gates of 0.25 ns."""
    g = np.random.default_rng(seed)
    c = (np.arange(n, dtype=np.float32) + 0.5) - n / 2.0
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    semi = np.array([0.40, 0.46, 0.44]) * n * g.uniform(0.95, 1.05, size=3)
    rr = np.sqrt((X / semi[0]) ** 2 + (Y / semi[1]) ** 2 + (Z / semi[2]) ** 2)
    rxy = np.sqrt(X * X + Y * Y)
    theta = np.arctan2(rxy, Z)
    phi = np.arctan2(Y, X)
    pert = _perturbed_radius(g, theta, phi, 0.04)
    rho = rr / pert  # < 1 inside the outer surface
    mean_semi = float(np.mean(semi))
    t1 = g.uniform(3, 5) / unitinmm / mean_semi
    t2 = g.uniform(5, 8) / unitinmm / mean_semi
    t3 = g.uniform(1, 3) / unitinmm / mean_semi
    vol = np.zeros((n, n, n), dtype=np.uint8)
    vol[rho < 1.0] = 1
    vol[rho < 1.0 - t1] = 2
    vol[rho < 1.0 - t1 - t2] = 3
    core = rho < 1.0 - t1 - t2 - t3
    split = (X * g.normal() + Y * g.normal() + Z * g.normal()) / n + 0.15 * _perturbed_radius(g, theta, phi, 1.0, 3) - 0.15
    vol[core & (split >= 0)] = 4
    vol[core & (split < 0)] = 5
    prop = np.array([[0, 0, 1, 1], [0.019, 7.8, 0.89, 1.37], [0.019, 7.8, 0.89, 1.37], [0.004, 0.009, 0.89, 1.37],
                     [0.02, 9.0, 0.89, 1.37], [0.08, 40.9, 0.89, 1.37]], dtype=np.float32)
    cfg = {
        "nphoton": int(nphoton),
        "vol": vol,
        "prop": prop,
        "unitinmm": unitinmm,
        "srcpos": [n / 2.0, n / 2.0, n - 0.5],
        "srcdir": [0, 0, -1],
        "issrcfrom0": 1,
        "tstart": 0.0,
        "tend": 5e-9,
        "tstep": 2.5e-10,
        "seed": seed,
        "isreflect": 1,
        "issavedet": 0,
    }
    cfg.update(kw)
    return cfg
