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
