"""CPU tests pinning the oracle (test infrastructure) before it judges the
HIP engine: RNG known answers, the host libc's rand(), elementary-function
accuracy, detector masking against the product's preprocessing, the
reference's absorbed-fraction known answers, energy conservation and
diffusion theory."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import diffusion
import oracle_py
from parity_util import oracle_run, seq_sum
from mcx_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "rng_kat.json")))


def test_glibc_rand_restatement_matches_host_libc():
    libc = ctypes.CDLL(None)
    for seed in (0, 1, 5, 12345, 1648335518, 0x623F9A9E):
        libc.srand(ctypes.c_uint(seed))
        want = [libc.rand() for _ in range(400)]
        assert list(oracle_py.glibc_rand(seed, 400)) == want, seed


def test_glibc_rand_default_sequence():
    # the well-known first outputs of glibc rand() with no srand()
    assert list(oracle_py.glibc_rand(0, 3)) == [1804289383, 846930886, 1681692777]


def test_rng_known_answer_stream_and_disk_order():
    seeds = oracle_py.seedwords(0, 4)
    got = oracle_py.rng_floats(seeds, 8)
    assert [float(x) for x in got] == KAT["stream_order"]
    disk = np.asarray(got)[KAT["disk_from_stream_permutation"]]
    assert [float(x) for x in disk] == KAT["disk_order"]


def test_xorshift_seed_word_packing_planted_error_detected():
    # a plausible mistake (rand() operands combined right to left) misses the KAT
    r = oracle_py.glibc_rand(0, 12).astype(np.uint32)
    wrong = [(int(r[3 * i + 2]) << 16 | int(r[3 * i + 1]) << 1 | int(r[3 * i]) >> 14) & 0xFFFFFFFF for i in range(4)]
    got = oracle_py.rng_floats(np.array(wrong, dtype=np.uint32), 8)
    assert [float(x) for x in got] != KAT["stream_order"]


def _ulp_err(f32, exact):
    f32 = np.asarray(f32, dtype=np.float64)
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(f32 - exact) / ulp


def test_expf_accuracy():
    L = oracle_py.lib()
    xs = np.concatenate([-np.logspace(-8, 1.9, 4000), np.linspace(-80, 0, 4001)]).astype(np.float32)
    got = np.array([L.mcx_oracle_expf(float(x)) for x in xs])
    assert _ulp_err(got, np.exp(xs.astype(np.float64))).max() <= 2.0
    assert L.mcx_oracle_expf(0.0) == 1.0 and L.mcx_oracle_expf(-200.0) == 0.0


def test_logf_accuracy():
    L = oracle_py.lib()
    xs = (np.arange(1, 8388609, 997, dtype=np.float64) / 8388608.0).astype(np.float32)
    xs = np.concatenate([xs + np.float32(1.1920929e-07), np.array([1.0000001, 0.5, 2.0 ** -23], np.float32)])
    got = np.array([L.mcx_oracle_logf(float(x)) for x in xs])
    exact = np.log(xs.astype(np.float64))
    nz = exact != 0
    assert _ulp_err(got[nz], exact[nz]).max() <= 2.0
    assert L.mcx_oracle_logf(1.0) == 0.0


def test_sincos2pi_accuracy():
    L = oracle_py.lib()
    s, c = ctypes.c_float(), ctypes.c_float()
    us = np.linspace(0, 1, 20001, dtype=np.float32)[:-1]
    err = 0.0
    for u in us:
        L.mcx_oracle_sincos2pi(float(u), ctypes.byref(s), ctypes.byref(c))
        a = 2 * math.pi * float(u)
        err = max(err, abs(s.value - math.sin(a)), abs(c.value - math.cos(a)))
        assert abs(s.value * s.value + c.value * c.value - 1) < 4e-7
    assert err < 3e-7


def test_maskdet_matches_product_preprocessing():
    """DET_MASK voxels of the product's mcx_maskdet (through the C ABI) equal
    the oracle's restatement (src/mcx_utils.c:4214-4340), bit for bit."""
    from mcx_amd import _lib
    L = _lib.lib()
    vol = workloads.sphere_volume(40, 12.0)
    vol[:, :, 30:] = 0
    det = np.array([[20, 20, 0, 2], [30, 20, 0, 3.5], [20, 20, 30, 2], [5, 5, 5, 1]], dtype=np.float32)
    want = oracle_py.maskdet(vol.astype(np.uint32), det)
    c = _lib.Config()
    L.mcx_initcfg(ctypes.byref(c))
    c.parentid = _lib.mpPython
    c.flog = None
    c.dim.x, c.dim.y, c.dim.z = vol.shape
    flat = np.asfortranarray(vol.astype(np.uint32)).ravel(order="F").copy()
    d4 = det.copy()
    c.vol = flat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))
    c.detpos = d4.ctypes.data_as(ctypes.POINTER(_lib.Float4))
    c.detnum = 4
    c.issavedet = 1
    L.mcx_maskdet(ctypes.byref(c))
    got = flat.reshape(vol.shape, order="F")
    assert np.array_equal(got, want)
    assert ((want & 0x80000000) != 0).sum() > 20


@pytest.mark.parametrize("reflect,lo,hi", [(0, 0.1592, 0.1946), (1, 0.27, 0.28)])
def test_cube60_absorbed_fraction_known_answer(reflect, lo, hi):
    """cube60 17.69% +-10% (speedcontest/mcxcontest:103, :262) without
    reflection; cube60b 27.x% (test/testmcx.sh:65) with Fresnel boundaries."""
    r = oracle_run(workloads.cube60(100000, reflect=reflect), 8192)
    assert r["errors"] == 0
    assert r["energytot"] == 100000.0
    assert lo <= r["absorbed"] <= hi, r["absorbed"]


def test_planted_absorption_error_leaves_known_answer_band():
    cfg = workloads.cube60(100000, reflect=0)
    cfg["prop"] = cfg["prop"].copy()
    cfg["prop"][1, 0] *= 1.5  # a mua scaling mistake
    r = oracle_run(cfg, 8192)
    assert not (0.1592 <= r["absorbed"] <= 0.1946)


def test_energy_conservation_with_tally():
    """absorbed energy = sum over voxels of mua * fluence (energy output) and
    launched = nphoton for a pencil beam (mcxlab/examples/mcx_gpu_benchmarks.m:56)."""
    cfg = workloads.cube60(50000)
    r = oracle_run(cfg, 4096, outputtype=2)
    tot = seq_sum(r["genergy"][1::2])
    esc = seq_sum(r["genergy"][0::2])
    assert tot == 50000.0
    dep = r["field"].sum()
    # the 0-5 ns window cuts a small tail, counted as escaped in both
    assert abs(dep - (tot - esc)) / (tot - esc) < 1e-5  # FP32 per-thread sums


def test_fp32_shadow_tally_close_to_fp64():
    r = oracle_run(workloads.cube60(50000), 4096, field32=True)
    a, b = r["field32"].astype(np.float64), r["field"]
    sig = b > 1e-3 * b.max()
    assert (np.abs(a[sig] - b[sig]) / b[sig]).max() < 1e-4


def test_homogeneous_fluence_matches_diffusion_theory():
    """CW fluence of cube60b vs the semi-infinite diffusion solution with
    Reff=0.493 for n=1.37 (mcxlab/examples/demo_validation_homogeneous.m:
    87-117), away from the source (>= 10 mm, ~10 transport lengths)."""
    r = oracle_run(workloads.cube60(200000), 16384)
    f = r["field"][..., 0] / r["energytot"]
    ks = np.arange(10, 21)
    det = np.stack([np.full(ks.shape, 29.5), np.full(ks.shape, 29.5), ks + 0.5], 1)
    cw = diffusion.cwdiffusion(0.005, 0.99, 0.493, [29, 29, 0], det)[:, 0]
    ratio = f[28:30, 28:30, ks].mean(axis=(0, 1)) / cw
    assert abs(ratio.mean() - 1) < 0.05 and np.all(np.abs(ratio - 1) < 0.15), ratio
    xs = np.arange(30, 42)
    det = np.stack([xs + 0.5, np.full(xs.shape, 29.5), np.full(xs.shape, 10.5)], 1)
    cw = diffusion.cwdiffusion(0.005, 0.99, 0.493, [29, 29, 0], det)[:, 0]
    ratio = 0.5 * (f[xs, 28, 10] + f[xs, 29, 10]) / cw
    assert abs(ratio.mean() - 1) < 0.05 and np.all(np.abs(ratio - 1) < 0.15), ratio


def test_tddiffusion_golden():
    """pmcx/test/test_utils.py:108-129"""
    cw = diffusion.tddiffusion(0.005, 1.0, 299792458000, 0, [29, 29, 0], np.array([[29, 13, 8], [0, 0, 0]]) + 1,
                               [2.5e-9])
    np.testing.assert_allclose(cw, [[1888.13933534], [92.31054244]], rtol=1e-9)


def test_cwdiffusion_is_time_integral_of_tddiffusion():
    t = np.linspace(1e-12, 2e-8, 200001)
    det = [[35.5, 29.5, 10.5]]
    td = diffusion.tddiffusion(0.005, 0.99, 299792458000 / 1.37, 0.493, [29, 29, 0], det, t)[0]
    cw = diffusion.cwdiffusion(0.005, 0.99, 0.493, [29, 29, 0], det)[0, 0]
    # cwdiffusion uses b = sqrt(3 mua musp) where the time integral gives
    # sqrt(3 mua (mua+musp)): a 0.26% difference at this distance
    assert abs(np.trapezoid(td, t) - cw) / cw < 5e-3


def test_detection_count_known_answer():
    """cube60b prints 'detected 4...' photons per 1e6 (test/testmcx.sh:81)."""
    det = np.array([[29, 19, 0, 1], [29, 39, 0, 1], [19, 29, 0, 1], [39, 29, 0, 1]], dtype=np.float32)
    r = oracle_run(workloads.cube60(200000, detpos=det, savedetflag="DP", issavedet=1), 8192)
    per_million = r["detected"] * 5
    assert 4000 <= per_million < 5000, per_million
    detid = r["detp"][:, 0]
    assert set(np.unique(detid)) <= {1, 2, 3, 4}


@pytest.mark.parametrize("bc,nphoton,lo,hi", [
    ("aarraa", 100000, 0.27, 0.28),   # test/testmcx.sh:73 (-b 0 -B aarraa: absorbed 27.x%)
    ("cccccc", 1000, 0.99, 1.0),      # test/testmcx.sh:77 (cyclic: absorbed 99.x%)
])
def test_boundary_condition_known_answers(bc, nphoton, lo, hi):
    cfg = workloads.cube60(nphoton, reflect=0, bc=bc)
    r = oracle_run(cfg, 2048)
    assert r["errors"] == 0
    esc, tot = seq_sum(r["genergy"][0::2]), seq_sum(r["genergy"][1::2])
    assert lo <= (tot - esc) / tot < hi


def test_boundary_detector_faces_known_answer():
    """test/testmcx.sh:112: with every face a detector ('______111111'),
    97-99 % of 1e4 photons are detected (the rest are cut by the 5 ns gate)"""
    cfg = workloads.cube60(10000, reflect=0, bc="______111111", issavedet=1, savedetflag="DP")
    r = oracle_run(cfg, 2048)
    assert 9700 <= r["detected"] <= 9999


@pytest.mark.parametrize("src", [
    dict(srctype="fourierx", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 40], srcparam2=[2, 1, 0.1, 0]),
    dict(srctype="fourierx2d", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 40], srcparam2=[2, 3, 0, 0.25]),
    dict(srctype="zgaussian", srcpos=[29.5, 29.5, 0], srcparam1=[0.3, 0, 0, 0]),
    dict(srctype="line", srcpos=[10, 29.5, 29.5], srcparam1=[40, 0, 0, 0]),
    dict(srctype="slit", srcpos=[10, 29.5, 0], srcparam1=[40, 0, 0, 0], srcparam2=[0.1, 0.05, 0, 0]),
    dict(srctype="hyperboloid", srcpos=[29.5, 29.5, 0], srcparam1=[1.0, 20.0, 5.0, 0]),
    dict(srctype="pattern", srcpos=[10, 10, 0], srcparam1=[40, 0, 0, 8], srcparam2=[0, 40, 0, 8],
         srcpattern=np.full((8, 8), 0.5, dtype=np.float32)),
])
def test_more_source_types_conserve_energy(src):
    """Every launched photon is accounted for (absorbed + escaped); patterned
    and modulated sources launch less than one unit of weight per photon
    (a uniform 0.5 pattern launches exactly 0.5), a line source inside the
    cube absorbs more than a surface source."""
    cfg = workloads.cube60(4000, reflect=0, **src)
    r = oracle_run(cfg, 512, want_field=False)
    assert r["errors"] == 0
    tot = seq_sum(r["genergy"][1::2])
    esc = seq_sum(r["genergy"][0::2])
    assert 0.0 < tot <= 4000.0 + 1e-6 and 0.0 <= esc <= tot
    if src["srctype"] == "pattern":
        assert tot == pytest.approx(2000.0, rel=1e-6)
    if src["srctype"] == "line":
        assert (tot - esc) / tot > 0.5


def test_diffuse_reflectance_tally():
    """issaveref (src/mcx_core.cu:1124-1145): the weight of photons leaving
    tissue into void voxels of the grid is tallied there with a negative
    sign; it is part of (and here most of) the escaped energy."""
    vol = np.ones((40, 40, 40), dtype=np.uint8)
    vol[:, :, :2] = 0
    cfg = workloads.cube60(20000, vol=vol, srcpos=[19.5, 19.5, 0], reflect=0, issaveref=1)
    r = oracle_run(cfg, 1024)
    f = r["field"][..., 0]
    void = vol == 0
    assert np.all(f[void] <= 0) and np.all(f[~void] >= 0)
    dref = -f[void].sum()
    esc = seq_sum(r["genergy"][0::2])
    assert 0.5 * esc < dref <= esc * (1 + 1e-9)


def _planar(**kw):
    """cube60planar (src/mcx_bench.c:158-205): a 40x40 planar source 10
    voxels above the cube, reflection on"""
    c = dict(srctype="planar", srcpos=[10.0, 10.0, -10.0], srcparam1=[40, 0, 0, 0], srcparam2=[0, 40, 0, 0], reflect=1)
    c.update(kw)
    return c


@pytest.mark.parametrize("src,band", [
    (_planar(), (25.0, 26.0)),                                                        # test/testmcx.sh:85
    (dict(srctype="isotropic", srcpos=[29, 29, 29], reflect=0), (88.0, 89.0)),        # :89
    (dict(srctype="cone", srcparam1=[0.5, 0, 0, 0], reflect=0,
          prop=np.array([[0, 0, 1, 1], [0.001, 0.001, 0, 1]], np.float32)), (6.0, 7.0)),  # :93
    (_planar(srctype="fourier", srcparam1=[40, 0, 0, 2]), (25.0, 26.0)),              # :97
    (_planar(srctype="pencilarray", srcparam1=[40, 0, 0, 4], srcparam2=[0, 20, 0, 2]), (23.0, 24.0)),  # :101
])
def test_source_type_known_answers(src, band):
    """Absorbed fractions the reference's test script greps for each source
    type ('absorbed:.*25\\.[0-9]+%' etc.), 2e5 photons."""
    src = dict(src)
    reflect = src.pop("reflect")
    cfg = workloads.cube60(200000, reflect=reflect, **src)
    r = oracle_run(cfg, 1024, want_field=False)
    assert r["errors"] == 0
    assert band[0] <= 100 * r["absorbed"] < band[1], r["absorbed"]


def test_2d_sheet_is_orientation_independent():
    """A 2-D domain (one axis of extent 1) scatters in its plane
    (rotatevector2d, src/mcx_core.cu:969-999).  The same sheet lying in the
    y-z plane (x singular) and in the x-y plane (z singular), with the
    source and direction rotated with it, is one problem: absorbed
    fractions agree within the photon noise and the fluence maps, mapped
    onto each other, agree."""
    n = 40000
    a = workloads.cube60(n, vol=np.ones((1, 40, 40), dtype=np.uint8), srcpos=[0, 19.5, 0], srcdir=[0, 0, 1])
    b = workloads.cube60(n, vol=np.ones((40, 40, 1), dtype=np.uint8), srcpos=[19.5, 0, 0], srcdir=[0, 1, 0])
    ra, rb = oracle_run(a, 4096), oracle_run(b, 4096)
    assert ra["errors"] == 0 and rb["errors"] == 0

    def absorbed(r):
        tot = seq_sum(r["genergy"][1::2])
        return (tot - seq_sum(r["genergy"][0::2])) / tot

    fa, fb = absorbed(ra), absorbed(rb)
    # in-plane paths: 0.18 absorbed; scattering in 3-D out of the 1-voxel
    # sheet (no 2-D mode) would absorb 0.02
    assert 0.1 < fa < 0.3, fa
    assert abs(fa - fb) < 5 * 0.5 / np.sqrt(n), (fa, fb)
    # sheet a: [1, y, z]; sheet b: [x, y, 1] with (x, y) of b = (y, z) of a
    ma = ra["field"][0, :, :, 0]
    mb = rb["field"][:, :, 0, 0]
    sig = ma > 0.02 * ma.max()
    rel = np.abs(ma[sig] - mb[sig]) / ma[sig]
    assert np.median(rel) < 0.05, np.median(rel)


def test_first_number_of_host_seeded_streams_is_biased():
    """Why lane refill keeps >= 64 photons per virtual thread by default:
    the host's seed words OR overlapping rand() values, so the first
    xorshift128+ output of each stream is far from uniform (the reference
    draws the same numbers); the second output is uniform."""
    n = 1 << 16
    w = oracle_py.seedwords(29012392, 4 * n).reshape(n, 4)
    u = np.array([oracle_py.rng_floats(w[i], 2) for i in range(n)], dtype=np.float64)
    e = n / 16.0
    chi = [float((((np.histogram(u[:, k], bins=16, range=(0, 1))[0] - e) ** 2) / e).sum()) for k in range(2)]
    assert chi[0] > 120, chi  # 15 degrees of freedom
    assert chi[1] < 50, chi
    assert np.mean(-np.log(u[:, 0] + 1e-30)) > 1.02  # first scattering length about 4 % long
