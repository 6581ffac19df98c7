"""Pins of the CPU oracle against what the mathematics and the reference's
published answers fix, each with a planted mistake that must fail it.

The oracle's transport loop calls the functions tested here (HG inversion,
face Fresnel, void-entry reflectance, refraction, voxel stepping, source
launch), so a pin of the function is a pin of the loop.  Mistakes are
planted in the oracle itself (oracle_py.mutation), never in the product.

Closed forms used:
  * Henyey-Greenstein: CDF F(mu) = (1-g^2)/(2g) [(1+g^2-2g mu)^-1/2 - (1+g)^-1]
    (integral of the HG pdf) and the Legendre moments E[P_l(mu)] = g^l;
  * Fresnel: the textbook s/p amplitude coefficients with Snell's law;
  * voxel stepping: exact rational ray/plane crossings;
  * sources: the distributions the reference documents for each type
    (mcxlab/mcxlab.m:190-250), restated as geometry or as the
    distribution of the documented sampling construction;
  * whole transport: the reference's published absorbed fractions
    (speedcontest/mcxcontest:103-105, mcxlab/examples/mcx_gpu_benchmarks.m:65-112)
    and a mirror face against the unfolded domain.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy import special, stats

import oracle_py
from mcx_amd import workloads
from parity_util import oracle_run

EPS = 1.1920928955078125e-07


# ---------------------------------------------------------------- HG sampling

def _hg_cdf(mu, g):
    return (1 - g * g) / (2 * g) * (1.0 / np.sqrt(1 + g * g - 2 * g * mu) - 1.0 / (1 + g))


def _uniforms(n, seed=2024):
    # the oracle's own xorshift128+ stream, as the loop feeds the inversion
    return oracle_py.rng_floats(oracle_py.seedwords(seed, 4), n)


def _hg_check(g, n=200_000):
    mu = oracle_py.hg_costheta(g, _uniforms(n)).astype(np.float64)
    ok = True
    # Legendre moments of HG: E[P_l(mu)] = g^l
    for l, P in ((1, mu), (2, 0.5 * (3 * mu * mu - 1)), (3, 0.5 * (5 * mu ** 3 - 3 * mu))):
        se = P.std() / np.sqrt(n)
        ok &= abs(P.mean() - g ** l) < 5 * se + 1e-6
    # chi-square over 50 equiprobable bins of the HG CDF
    edges = np.interp(np.linspace(0, 1, 51), _hg_cdf(np.linspace(-1, 1, 20001), g), np.linspace(-1, 1, 20001))
    edges[0], edges[-1] = -1.0 - 1e-9, 1.0 + 1e-9
    counts, _ = np.histogram(mu, edges)
    chi2 = ((counts - n / 50) ** 2 / (n / 50)).sum()
    ok &= stats.chi2.sf(chi2, 49) > 1e-4
    return bool(ok)


@pytest.mark.parametrize("g", [0.01, 0.5, 0.9, 0.95, -0.5])
def test_hg_sampling_matches_hg_distribution(g):
    assert _hg_check(g)


def test_hg_isotropic_branch():
    mu = oracle_py.hg_costheta(1e-8, _uniforms(200_000)).astype(np.float64)
    assert stats.kstest(mu, "uniform", args=(-1, 2)).pvalue > 1e-3


@pytest.mark.parametrize("mut", [1, 2])
def test_hg_planted_mistakes_fail(mut):
    with oracle_py.mutation(mut):
        assert not _hg_check(0.9)


# -------------------------------------------------------------------- Fresnel

def _fresnel_textbook(n1, n2, ci):
    st2 = (n1 / n2) ** 2 * (1 - ci * ci)
    if st2 >= 1:
        return 1.0
    ct = np.sqrt(1 - st2)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci)
    return 0.5 * (rs * rs + rp * rp)


NPAIRS = [(1.37, 1.0), (1.0, 1.37), (1.33, 1.45), (1.45, 1.33), (1.37, 1.37)]


def _fresnel_check(fn):
    for n1, n2 in NPAIRS:
        for ci in np.linspace(0.02, 1.0, 99):
            ci = float(np.float32(ci))
            want = _fresnel_textbook(n1, n2, ci)
            if abs(fn(n1, n2, ci) - want) > 2e-6 + 2e-5 * want:
                return False
    return True


def test_face_fresnel_matches_textbook():
    assert _fresnel_check(oracle_py.fresnel_face)


def test_void_entry_reflectance_matches_textbook():
    assert _fresnel_check(oracle_py.reflectcoeff)


def test_fresnel_special_cases():
    for fn in (oracle_py.fresnel_face, oracle_py.reflectcoeff):
        # normal incidence
        assert abs(fn(1.0, 1.37, 1.0) - (0.37 / 2.37) ** 2) < 1e-7
        # total internal reflection beyond the critical angle of 1.37 -> 1
        crit = np.sqrt(1 - (1.0 / 1.37) ** 2)
        assert fn(1.37, 1.0, float(crit) * 0.99) == 1.0
        assert fn(1.37, 1.0, float(crit) * 1.01) < 1.0
        # Brewster angle: the p-term vanishes, R = rs^2 / 2
        tb = np.arctan(1.37 / 1.0)
        ci = float(np.cos(tb))
        ct = np.sqrt(1 - (1.0 / 1.37) ** 2 * (1 - ci * ci))
        rs = (1.0 * ci - 1.37 * ct) / (1.0 * ci + 1.37 * ct)
        assert abs(fn(1.0, 1.37, ci) - 0.5 * rs * rs) < 2e-6


@pytest.mark.parametrize("mut,fn", [(3, oracle_py.fresnel_face), (6, oracle_py.reflectcoeff)])
def test_fresnel_planted_mistakes_fail(mut, fn):
    with oracle_py.mutation(mut):
        assert not _fresnel_check(fn)


def _snell_check():
    rng = np.random.default_rng(5)
    for _ in range(300):
        face = int(rng.integers(0, 3))
        n1, n2 = NPAIRS[int(rng.integers(0, 4))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        ci = abs(v[face])
        if (n1 / n2) ** 2 * (1 - ci * ci) >= 0.98:  # at or near total internal reflection
            continue
        out = oracle_py.transmit(n1, n2, face, v.astype(np.float32)).astype(np.float64)
        tan_in = np.delete(v, face)
        tan_out = np.delete(out, face)
        # tangential component scaled by n1/n2, unit length, same side
        if np.abs(tan_out - n1 / n2 * tan_in).max() > 2e-6:
            return False
        if abs(np.linalg.norm(out) - 1) > 2e-6 or np.sign(out[face]) != np.sign(v[face]):
            return False
    return True


def test_transmit_obeys_snell():
    assert _snell_check()


def test_transmit_planted_mistake_fails():
    with oracle_py.mutation(5):
        assert not _snell_check()


# -------------------------------------------------------------- voxel stepping

def _exact_trace(p, v, dims, maxn=4096):
    """Voxels a ray enters, by exact rational crossings; also the smallest
    relative gap between competing crossings (near-ties are ambiguous in
    binary32)."""
    p = [Fraction(float(x)) for x in p]
    v = [Fraction(float(x)) for x in v]
    cell = [int(np.floor(float(x))) for x in p]
    t = Fraction(0)
    seq, times, gap = [], [], np.inf
    for _ in range(maxn):
        ts = []
        for a in range(3):
            if v[a] > 0:
                ts.append((cell[a] + 1 - p[a]) / v[a])
            elif v[a] < 0:
                ts.append((cell[a] - p[a]) / v[a])
            else:
                ts.append(None)
        cand = sorted((x, a) for a, x in enumerate(ts) if x is not None)
        tn, ax = cand[0]
        if len(cand) > 1:
            gap = min(gap, float(cand[1][0] - tn))
        cell[ax] += 1 if v[ax] > 0 else -1
        t = tn
        seq.append(tuple(cell))
        times.append(float(t))
        if not all(0 <= cell[a] < dims[a] for a in range(3)):
            break
    return seq, np.array(times), gap


def _trace_check(nray=200):
    rng = np.random.default_rng(11)
    dims = (7, 5, 6)
    compared = 0
    for i in range(nray):
        p = (rng.uniform(0, 1, 3) * np.array(dims)).astype(np.float32)
        if i < 6:  # axis-aligned rays
            v = np.zeros(3, np.float32)
            v[i % 3] = 1.0 if i < 3 else -1.0
        else:
            v = rng.normal(size=3)
            v = (v / np.linalg.norm(v)).astype(np.float32)
        ids, lens = oracle_py.trace(p, v, dims)
        seq, times, gap = _exact_trace(p, v, dims)
        if gap < 1e-4:
            continue
        compared += 1
        if [tuple(x) for x in ids] != seq:
            return False
        if np.abs(np.cumsum(lens.astype(np.float64)) - times).max() > 1e-4:
            return False
    assert compared > nray * 0.8
    return True


def test_voxel_stepping_matches_exact_crossings():
    assert _trace_check()


def test_hitgrid_planted_mistake_fails():
    with oracle_py.mutation(4):
        assert not _trace_check()


# ------------------------------------------------------------------- sources

N_SRC = 20_000
PMIN = 1e-3


def _frame(d):
    d = np.asarray(d, dtype=np.float64)
    d = d / np.linalg.norm(d)
    a = np.array([1.0, 0, 0]) if abs(d[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(d, a)
    e1 /= np.linalg.norm(e1)
    return d, e1, np.cross(d, e1)


def _fourierx_ok(kind):
    S = np.array([10.0, 12.0, 30.0])
    d = np.array([0.0, 0.0, 1.0])
    v1 = np.array([30.0, 10.0, 0.0])
    v2n = 20.0
    s2 = [2, 1, 0.1, 0.3] if kind == "fourierx" else [2, 3, 0.1, 0.25]
    p, w, _ = oracle_py.launch_samples(kind, S, d, N_SRC, srcparam1=list(v1) + [v2n], srcparam2=s2)
    # documented: srcdir x v1 = v2 (normalised), |v2| = srcparam1(4)
    v2 = v2n * np.cross(d, v1 / np.linalg.norm(v1))
    A = np.stack([v1[:2], v2[:2]], axis=1)
    xy = np.linalg.solve(A, (p[:, :2] - S[:2]).T).T
    x, y = xy[:, 0], xy[:, 1]
    if np.abs(p[:, 2] - S[2]).max() > 1e-4:
        return False
    if x.min() < -1e-5 or x.max() > 1 + 1e-5 or y.min() < -1e-5 or y.max() > 1 + 1e-5:
        return False
    if stats.kstest(x, "uniform").pvalue < PMIN or stats.kstest(y, "uniform").pvalue < PMIN:
        return False
    kx, ky = s2[0], s2[1]
    if kind == "fourierx":
        # S = 0.5 [1 + M' cos(2 pi (kx x + ky y) + 2 pi phi0)], modulation
        # depth M' = 1 - srcparam2(4) as documented for the fourier source
        want = 0.5 * (1 + (1 - s2[3]) * np.cos(2 * np.pi * (kx * x + ky * y) + 2 * np.pi * s2[2]))
    else:
        # separable 2-D basis with phases phix, phiy (x 2 pi)
        want = 0.5 * (1 + np.sin(2 * np.pi * kx * x + 2 * np.pi * s2[2]) * np.sin(2 * np.pi * ky * y + 2 * np.pi * s2[3]))
    if np.abs(w - want).max() > 5e-5:
        return False
    # integer frequencies: the pattern averages to one half
    return abs(w.mean() - 0.5) < 5 * w.std() / np.sqrt(len(w))


@pytest.mark.parametrize("kind,mut", [("fourierx", 11), ("fourierx2d", 12)])
def test_fourier_sources(kind, mut):
    assert _fourierx_ok(kind)
    with oracle_py.mutation(mut):
        assert not _fourierx_ok(kind)


def _zgaussian_ok():
    sigma = 0.2
    d = [0.0, 0.0, 1.0]
    _, _, v = oracle_py.launch_samples("zgaussian", [32, 32, 32], d, N_SRC, srcparam1=[sigma, 0, 0, 0])
    ang = np.arccos(np.clip(v[:, 2], -1, 1))
    # theta = sigma * R * U with R ~ Rayleigh(1) (sqrt(2E), E ~ Exp(1)) and
    # U ~ U(-1,1): P(|R U| <= t) = sqrt(pi/2) t erfc(t/sqrt 2) + 1 - exp(-t^2/2)
    def cdf(a):
        t = np.asarray(a) / sigma
        return np.sqrt(np.pi / 2) * t * special.erfc(t / np.sqrt(2)) + 1 - np.exp(-t * t / 2)
    if stats.kstest(ang, cdf).pvalue < PMIN:
        return False
    phi = np.arctan2(v[:, 1], v[:, 0])
    return stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue >= PMIN


def test_zgaussian_source():
    assert _zgaussian_ok()
    with oracle_py.mutation(13):
        assert not _zgaussian_ok()


def _line_ok(fan):
    S = np.array([10.0, 29.5, 29.5])
    s1 = np.array([40.0, 0.0, 10.0])
    d = np.array([0.0, 0.6, 0.8])
    p, _, v = oracle_py.launch_samples("line", S, d, N_SRC, srcparam1=list(s1) + [0], srcparam2=[fan, 0, 0, 0])
    a = s1 / np.linalg.norm(s1)
    rel = p - S
    t = rel @ s1 / (s1 @ s1)
    if np.abs(rel - t[:, None] * s1).max() > 1e-4 or stats.kstest(t, "uniform").pvalue < PMIN:
        return False
    # directions perpendicular to the line
    if np.abs(v @ a).max() > 1e-5:
        return False
    dperp = d - (d @ a) * a
    dperp /= np.linalg.norm(dperp)
    nrm = np.cross(a, dperp)
    psi = np.arctan2(v @ nrm, v @ dperp)
    if fan == 0:
        # uniform around the line
        return stats.kstest(psi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue >= PMIN
    # fan: rotated within +-fan from the plane of srcdir and the line
    return stats.kstest(psi, "uniform", args=(-fan, 2 * fan)).pvalue >= PMIN


def test_line_source():
    assert _line_ok(0.0)
    assert _line_ok(0.5)
    with oracle_py.mutation(14):
        assert not _line_ok(0.5)


def _slit_ok():
    S = np.array([10.0, 29.5, 0.0])
    s1 = np.array([40.0, 0.0, 0.0])
    d = np.array([0.0, 0.0, 1.0])
    wx, wy = 0.1, 0.05
    p, _, v = oracle_py.launch_samples("slit", S, d, N_SRC, srcparam1=list(s1) + [0], srcparam2=[wx, wy, 0, 0])
    t = (p - S) @ s1 / (s1 @ s1)
    if stats.kstest(t, "uniform").pvalue < PMIN:
        return False
    # broadening: Gaussian of width srcparam2.x along s1 x srcdir and of
    # width srcparam2.y along the slit, in the plane normal to srcdir
    c = np.cross(s1, d)
    c /= np.linalg.norm(c)
    a = s1 / np.linalg.norm(s1)
    bx = (v @ c) / (v @ d)
    by = (v @ a) / (v @ d)
    return (stats.kstest(bx, "norm", args=(0, wx)).pvalue >= PMIN and
            stats.kstest(by, "norm", args=(0, wy)).pvalue >= PMIN)


def test_slit_source():
    assert _slit_ok()
    with oracle_py.mutation(15):
        assert not _slit_ok()


def _hyperboloid_ok(d):
    S = np.array([32.0, 32.0, 20.0])
    w0, zf, zr = 3.0, 8.0, 5.0
    d, _, _ = _frame(d)
    p, _, v = oracle_py.launch_samples("hyperboloid", S, d, N_SRC, srcparam1=[w0, zf, zr, 0])
    rel = p - S
    # launched in the plane through srcpos normal to srcdir
    if np.abs(rel @ d).max() > 2e-4:
        return False
    def radial(q):
        return np.linalg.norm(q - (q @ d)[:, None] * d[None, :], axis=1)
    # each ray is a straight line whose distance to the axis is smallest at
    # the focus zf along srcdir and grows as r(z) = r0 sqrt(1 + ((z-zf)/zr)^2)
    tf = zf / (v @ d)
    r_focus = radial(rel + tf[:, None] * v)
    r_launch = radial(rel)
    if np.abs(r_launch - r_focus * np.sqrt(1 + (zf / zr) ** 2)).max() > 2e-3 * (1 + r_launch.max()):
        return False
    r_beyond = radial(rel + (tf * 1.5)[:, None] * v)
    if np.abs(r_beyond - r_focus * np.sqrt(1 + (0.5 * zf / zr) ** 2)).max() > 2e-3 * (1 + r_launch.max()):
        return False
    # Gaussian beam: intensity exp(-2 r^2 / w0^2) at the waist, i.e. r^2 ~ Exp(mean w0^2 / 2)
    return stats.kstest(r_focus ** 2, "expon", args=(0, w0 * w0 / 2)).pvalue >= PMIN


@pytest.mark.parametrize("d", [[0, 0, 1], [0.3, 0.2, 0.93]])
def test_hyperboloid_source(d):
    assert _hyperboloid_ok(d)
    with oracle_py.mutation(16):
        assert not _hyperboloid_ok(d)


def _pattern3d_ok():
    S = np.array([20.0, 22.0, 24.0])
    dims = (4, 3, 5)
    pat = np.random.default_rng(3).uniform(0.05, 1.0, size=dims).astype(np.float32)
    p, w, _ = oracle_py.launch_samples("pattern3d", S, [0, 0, 1], N_SRC, srcparam1=list(dims) + [0], srcpattern=pat)
    rel = p - S
    for k in range(3):
        if rel[:, k].min() < 0 or rel[:, k].max() > dims[k]:
            return False
        if stats.kstest(rel[:, k] / dims[k], "uniform").pvalue < PMIN:
            return False
    # weight = the pattern cell containing the launch point
    idx = np.minimum(np.floor(rel).astype(int), np.array(dims) - 1)
    want = pat[idx[:, 0], idx[:, 1], idx[:, 2]]
    return np.mean(np.abs(w - want) > 1e-6) < 2e-3


def test_pattern3d_source():
    assert _pattern3d_ok()
    with oracle_py.mutation(17):
        assert not _pattern3d_ok()


def _arcsine_ok():
    d, e1, e2 = _frame([0.2, -0.3, 0.93])
    _, _, v = oracle_py.launch_samples("arcsine", [32, 32, 32], d, N_SRC)
    # zenith angle uniform on [0, pi], azimuth uniform
    th = np.arccos(np.clip(v @ d, -1, 1))
    phi = np.arctan2(v @ e2, v @ e1)
    return (stats.kstest(th, "uniform", args=(0, np.pi)).pvalue >= PMIN and
            stats.kstest(phi, "uniform", args=(-np.pi, 2 * np.pi)).pvalue >= PMIN)


def test_arcsine_source():
    assert _arcsine_ok()
    with oracle_py.mutation(18):
        assert not _arcsine_ok()


def _ring_ok(sector):
    S = np.array([32.0, 32.0, 10.0])
    r1, r2 = 10.0, 4.0
    s1 = [r1, r2, 0.5, 2.0] if sector else [r1, r2, 0, 0]
    p, _, v = oracle_py.launch_samples("ring", S, [0, 0, 1], N_SRC, srcparam1=s1)
    if np.abs(v - np.array([0, 0, 1.0])).max() > 1e-7 or np.abs(p[:, 2] - S[2]).max() > 1e-5:
        return False
    rel = p[:, :2] - S[:2]
    r2s = (rel ** 2).sum(1)
    # uniform over the annulus area: r^2 uniform in [r2^2, r1^2]
    if stats.kstest(r2s, "uniform", args=(r2 * r2, r1 * r1 - r2 * r2)).pvalue < PMIN:
        return False
    phi = np.mod(np.arctan2(rel[:, 1], rel[:, 0]), 2 * np.pi)
    lo, hi = (0.5, 2.0) if sector else (0.0, 2 * np.pi)
    return stats.kstest(phi, "uniform", args=(lo, hi - lo)).pvalue >= PMIN


@pytest.mark.parametrize("sector", [False, True])
def test_ring_source(sector):
    assert _ring_ok(sector)
    with oracle_py.mutation(19):
        assert not _ring_ok(sector)


# --------------------------------------------- published whole-transport answers

NTHREAD = 16384


def _absorbed_and_se(r):
    g = r["genergy"].astype(np.float64)
    esc, tot = g[0::2], g[1::2]
    a = (tot - esc) / tot
    return r["absorbed"], a.std(ddof=1) / np.sqrt(len(a))


def _contest_ok(name, nphoton):
    r = oracle_run(workloads.contest(name, nphoton), NTHREAD)
    assert r["errors"] == 0
    a, se = _absorbed_and_se(r)
    # 4 standard errors of this sample plus the rounding of the published 4 digits
    return abs(a - workloads.CONTEST_ABSORBED[name]) <= 4 * se + 5e-5, a, se


@pytest.mark.parametrize("name", ["cube60", "cubesph60b", "cube60planar"])
def test_contest_absorbed_fraction(name):
    ok, a, se = _contest_ok(name, 3_000_000)
    assert ok, (name, a, se)
    assert 4 * se + 5e-5 < 5e-3 * workloads.CONTEST_ABSORBED[name]


@pytest.mark.parametrize("name,mut", [("cubesph60b", 1), ("cubesph60b", 3), ("cube60planar", 3), ("cube60", 4)])
def test_contest_planted_mistakes_fail(name, mut):
    with oracle_py.mutation(mut):
        ok, a, se = _contest_ok(name, 1_000_000)
    assert not ok, (name, mut, a, se)


# ----------------------------------------------------- mirror face, unfolded

def _slab_profile(r):
    return r["field"][..., 0].sum(axis=(0, 1))


def test_mirror_face_equals_unfolded_domain():
    """A mirror at z=30 of a 60x60x30 cube is the 60x60x60 cube folded at
    z=30: same absorbed fraction, and each slab's fluence is the sum of
    its two images."""
    half = workloads.cube60(400_000, reflect=1, vol=np.ones((60, 60, 30), np.uint8), bc="rrrrrm")
    full = workloads.cube60(400_000, reflect=1, bc="rrrrrr")
    rh = oracle_run(half, 8192)
    rf = oracle_run(full, 8192)
    ah, seh = _absorbed_and_se(rh)
    af, sef = _absorbed_and_se(rf)
    assert abs(ah - af) < 4 * np.hypot(seh, sef), (ah, af)
    ph = _slab_profile(rh)
    pf = _slab_profile(rf)
    folded = pf[:30] + pf[::-1][:30]
    assert np.abs(ph / folded - 1).max() < 0.03
    # the same fold against an absorbing face: a mirror mistaken for 'a'
    ra = oracle_run(dict(half, bc="rrrrra"), 8192)
    aa, sea = _absorbed_and_se(ra)
    assert abs(aa - af) > 10 * np.hypot(sea, sef)
    assert np.abs(_slab_profile(ra) / folded - 1).max() > 0.2
