"""Property-word media formats (src/mcx_const.h:72-80).

Encoders (mcx_amd.media, the library's mcx_float2half2) are pinned by words
worked out by hand from the reference encoder's bit rules
(src/mcx_utils.c:5490-5546, src/pmcx.cpp:102-380) and by the error bounds
those rules imply.  The oracle's decoding (src/mcx_core.cu:582-687) is
pinned by equivalences: a word volume that encodes exactly the properties
of a label volume must reproduce the label run bit for bit, and a decoding
that swaps, drops or mis-scales a component breaks the equality.
"""
import ctypes

import numpy as np
import pytest

import mcx_amd
from mcx_amd import _lib, media
from parity_util import oracle_run

NT = 256


def _f2h_c(a, b):
    arr = (ctypes.c_float * 2)(a, b)
    return _lib.lib().mcx_float2half2(arr) & 0xFFFFFFFF


def _half(h):
    return float(np.array([h], dtype=np.uint16).view(np.float16)[0])


# ---- encoders ------------------------------------------------------------

# (mua, mus) -> word, worked by hand:
#  0.5: exponent 126 -> half exponent 14 (0b01110), mantissa 0; the first
#       value's low 3 mantissa bits are read back from the half just
#       written: bits 13-15 of 0x3800 = 0b001 -> 0x3801.  0.25 -> 0x3400.
#  1.0: exponent 15 (0b01111) -> 0x3C00, low bits 0b001 -> 0x3C01.
#  2.0: exponent 16 (0b10000) -> 0x4000, low bits from 0x4000 >> 13 = 0b010.
#  -0.5: 0xB800, bits 13-15 = 0b101 -> 0xB805; mus 0 -> denormal path -> 0.
#  (0, 0): packed 0 -> 0x00010000 (smallest mus, keeps the voxel non-void).
#  mus 1 + 2^-10 + 2^-11: mantissa truncated -> 0x3C01 (not rounded to 0x3C02).
#  mus 2^-20: denormal path, (0x800 >> 7) + ((0x800 >> 6) & 1) = 16 -> 0x0010.
HAND_WORDS = [
    ((0.5, 0.25), 0x34003801),
    ((1.0, 1.0), 0x3C003C01),
    ((2.0, 1.0), 0x3C004002),
    ((-0.5, 0.0), 0x0000B805),
    ((0.0, 0.0), 0x00010000),
    ((0.5, 1.0 + 2.0 ** -10 + 2.0 ** -11), 0x3C013801),
    ((0.5, 2.0 ** -20), 0x00103801),
]


@pytest.mark.parametrize("ab,word", HAND_WORDS)
def test_float2half2_hand_words(ab, word):
    assert int(media.float2half2(np.float32(ab[0]), np.float32(ab[1])).item()) == word
    assert _f2h_c(*ab) == word


def test_float2half2_c_and_numpy_agree():
    rng = np.random.default_rng(5)
    a = (10.0 ** rng.uniform(-9, 4.5, 4000) * rng.choice([-1, 1], 4000)).astype(np.float32)
    b = (10.0 ** rng.uniform(-9, 4.5, 4000)).astype(np.float32)
    a[:5] = [0.0, -0.0, 1.0, 65504.0, 6.1e-5]
    w = media.float2half2(a, b)
    for i in range(0, 4000, 7):
        assert int(w[i]) == _f2h_c(float(a[i]), float(b[i])), (a[i], b[i])


def test_float2half2_error_bounds():
    """Normal range: the second value is truncated (below the value by less
    than one half ulp, 2^-10 relative); the first value's three low bits are
    replaced, so it is within 2^-7 relative."""
    rng = np.random.default_rng(6)
    a = (10.0 ** rng.uniform(-4, 4, 2000)).astype(np.float32)
    b = (10.0 ** rng.uniform(-4, 4, 2000)).astype(np.float32)
    w = media.float2half2(a, b)
    lo = (w & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
    hi = (w >> 16).astype(np.uint16).view(np.float16).astype(np.float64)
    assert np.all(hi <= b) and np.all(b - hi < b * 2.0 ** -10 * 1.0001)
    assert np.all(np.abs(lo - a) < a * 2.0 ** -7 * 1.0001)


def test_half_truncate_exact_values():
    v = np.array([0.5, 0.75, 3.0, 1.375, 0.0078125, 100.0], dtype=np.float32)
    assert np.array_equal(media.half_truncate(v).astype(np.uint16), v.astype(np.float16).view(np.uint16))
    # 0.1 truncates (float16(0.1) rounds up to 0x2E66; truncation keeps 0x2E66 - 0 or 1 below)
    t = int(np.asarray(media.half_truncate(np.float32(0.1))).item())
    assert _half(t) <= 0.1 < _half(t + 1)


def test_encode_volume_packings():
    rng = np.random.default_rng(7)
    b = rng.integers(0, 128, size=(4, 3, 2, 2)).astype(np.int8)
    w, mb, dims = media.encode_volume(b)
    assert mb == 103 and dims == (3, 2, 2)
    u = b.astype(np.uint8).astype(np.uint32)
    assert w[1] == u[0, 1, 0, 0] | (u[1, 1, 0, 0] << 8) | (u[2, 1, 0, 0] << 16) | (u[3, 1, 0, 0] << 24)
    s = np.zeros((3, 2, 2, 2), dtype=np.int16)
    s[:, 1, 1, 1] = [3, 2, 40000 - 65536]
    w, mb, _ = media.encode_volume(s)
    assert mb == 98 and w[7] == 3 | (2 << 8) | ((40000 & 0x7FFF) << 16)
    s2 = np.zeros((2, 2, 2, 2), dtype=np.uint16)
    s2[:, 0, 1, 0] = [5, 65535]
    w, mb, _ = media.encode_volume(s2)
    assert mb == 104 and w[2] == 5 | (65535 << 16)
    f = np.zeros((3, 2, 1, 1), dtype=np.float32)
    f[:, 1, 0, 0] = [1.5, 2, 7]   # value 1.5 for property 2 (g) of label 7
    w, mb, _ = media.encode_volume(f, unitinmm=0.5)
    # index 2 is scaled by unitinmm in the reference's encoder: 0.75 -> 0x3A00
    assert mb == 99 and w[1] == (2 << 14 | 7) | (0x3A00 << 16)
    m = np.array([[[[0.0, np.nan, 0.25]]]], dtype=np.float32).reshape(1, 3, 1, 1)
    w, mb, _ = media.encode_volume(m, unitinmm=2.0)
    assert mb == 101
    assert list(w) == [np.float32(np.finfo(np.float32).eps).view(np.uint32), 0, np.float32(0.5).view(np.uint32)]
    a = np.zeros((4, 1, 1, 2), dtype=np.float32)
    a[:, 0, 0, 0] = [0.5, 0.25, 0.5, 1.0]
    a[:, 0, 0, 1] = [np.nan, 1.0, 0.5, 1.0]
    w, mb, dims = media.encode_volume(a)
    assert mb == 96 and w.size == 4
    assert list(w) == [0x34003801, 0, 0x3C003801, 0x3C003801]
    with pytest.raises(ValueError):
        media.encode_volume(np.full((3, 2, 2, 2), 5.0, dtype=np.float32))


# ---- oracle equivalences ---------------------------------------------------

def _base(prop, vol, nphoton=4000, **kw):
    cfg = {
        "nphoton": nphoton, "vol": vol, "prop": np.asarray(prop, dtype=np.float32),
        "srcpos": [10.5, 10.5, 0], "srcdir": [0, 0, 1], "issrcfrom0": 1,
        "tstart": 0.0, "tend": 5e-9, "tstep": 1e-9, "seed": 1234, "isreflect": 1, "issavedet": 0,
    }
    cfg.update(kw)
    return cfg


def _labels():
    v = np.ones((21, 21, 21), dtype=np.uint8)
    v[:3] = 0          # void slab: reflection into void and relaunch
    return v


def _same(a, b):
    assert np.array_equal(a["genergy"], b["genergy"])
    assert np.array_equal(a["field"], b["field"])


def test_mua_float_equals_label_volume():
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37]]
    v = _labels()
    ref = oracle_run(_base(prop, v), NT)
    m = np.where(v == 1, np.float32(0.01), np.float32(np.nan))[None].astype(np.float32)
    got = oracle_run(_base(prop, m), NT)
    _same(ref, got)
    # a different per-voxel mua changes the result
    m2 = m.copy()
    m2[0, 10:, :, 5:] = 0.02
    assert not np.array_equal(oracle_run(_base(prop, m2), NT)["field"], ref["field"])


def test_muamus_half_equals_label_volume():
    prop = [[0, 0, 1, 1], [0.0078125, 1.0, 0.9, 1.37]]
    v = _labels()
    ref = oracle_run(_base(prop, v), NT)
    w = np.where(v == 1, media.half_words(0.0078125, 1.0)[()], 0).astype(np.uint32)
    _same(ref, oracle_run(_base(prop, w, mediabyte=102), NT))
    # mua and mus swapped in the word: different result
    ws = np.where(v == 1, media.half_words(1.0, 0.0078125)[()], 0).astype(np.uint32)
    assert not np.array_equal(oracle_run(_base(prop, ws, mediabyte=102), NT)["genergy"], ref["genergy"])


def test_asgn_byte_equals_label_volume():
    # rows 1 and 2 differ in mua, mus, g and share n: bytes 0 for mua/mus/g
    # and any n byte give row 1 exactly.  Reflection is off: this format
    # decodes a void word to row 1 as well, so its boundaries with the void
    # are index-matched where a label volume's are not.
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.05, 3.0, 0.1, 1.37]]
    v = _labels()
    ref = oracle_run(_base(prop, v, isreflect=0), NT)
    b = np.zeros((4,) + v.shape, dtype=np.int8)
    b[3][v == 1] = 77
    _same(ref, oracle_run(_base(prop, b, isreflect=0), NT))
    b2 = b.copy()
    b2[0][v == 1] = 77          # mua byte set: mua moves towards row 2
    assert not np.array_equal(oracle_run(_base(prop, b2, isreflect=0), NT)["genergy"], ref["genergy"])


def test_asgn_byte_void_boundary_is_index_matched():
    """With reflection on, leaving tissue of row 1's index for the void never
    reflects (the void word decodes to row 1), so nothing returns from the
    void boundary: absorption is below the label run's, which reflects part
    of the light back at the 1.37/1.0 interface."""
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.05, 3.0, 0.1, 1.37]]
    v = _labels()
    b = np.zeros((4,) + v.shape, dtype=np.int8)
    b[3][v == 1] = 77
    lab = oracle_run(_base(prop, v, nphoton=20000), NT)
    byt = oracle_run(_base(prop, b, nphoton=20000), NT)
    noref = oracle_run(_base(prop, v, nphoton=20000, isreflect=0), NT)
    assert byt["absorbed"] < lab["absorbed"] - 0.02
    assert abs(byt["absorbed"] - noref["absorbed"]) < 0.01


def test_muamus_short_equals_label_volume():
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.05, 1.0, 0.9, 1.37]]
    v = _labels()
    ref = oracle_run(_base(prop, v), NT)
    s = np.zeros((2,) + v.shape, dtype=np.uint16)
    s[1][v == 1] = 1234         # mus fraction, rows share mus
    _same(ref, oracle_run(_base(prop, s), NT))
    s2 = s[::-1].copy()         # fraction moved to mua
    assert not np.array_equal(oracle_run(_base(prop, s2), NT)["genergy"], ref["genergy"])


def test_labelplus_equals_label_volume():
    prop = [[0, 0, 1, 1], [0.0078125, 1.0, 0.9, 1.375], [0.03, 2.0, 0.5, 1.5]]
    v = _labels()
    v[:, :, 12:] = 2
    ref = oracle_run(_base(prop, v), NT)
    f = np.zeros((3,) + v.shape, dtype=np.float32)
    f[2] = v
    f[1][v == 1] = 0            # replace mua of label 1 by its own value
    f[0][v == 1] = 0.0078125
    f[1][v == 2] = 3            # replace n of label 2 by its own value
    f[0][v == 2] = 1.5
    _same(ref, oracle_run(_base(prop, f), NT))
    f2 = f.copy()
    f2[0][v == 2] = 1.25        # a different n for label 2
    assert not np.array_equal(oracle_run(_base(prop, f2), NT)["genergy"], ref["genergy"])


def test_asgn_float_equals_label_volume():
    # values whose encoding is exact (the first component of each word
    # needs a half exponent below 8, see HAND_WORDS)
    row = [0.005859375, 1.0, 0.001953125, 1.375]
    prop = [[0, 0, 1, 1], row]
    v = _labels()
    ref = oracle_run(_base(prop, v), NT)
    a = np.zeros((4,) + v.shape, dtype=np.float32)
    for k in range(4):
        a[k][v == 1] = row[k]
    a[0][v == 0] = np.nan
    a[2][v == 0] = 0.001953125
    a[3][v == 0] = 1.0
    _same(ref, oracle_run(_base(prop, a), NT))
    a2 = a.copy()
    a2[3][v == 1] = 1.5         # per-voxel n differs from the label's
    assert not np.array_equal(oracle_run(_base(prop, a2), NT)["genergy"], ref["genergy"])


def test_mixlabel_zero_fraction_equals_label_volume():
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.05, 2.0, 0.5, 1.5]]
    v = _labels()
    ref = oracle_run(_base(prop, v), NT)
    s = np.zeros((3,) + v.shape, dtype=np.int16)
    s[0] = v
    s[1][v == 1] = 2
    _same(ref, oracle_run(_base(prop, s), NT))
    # with a fraction the second label is drawn: absorption moves towards label 2
    ref2 = oracle_run(_base(prop, np.where(v == 1, 2, 0).astype(np.uint8)), NT)
    s[2][v == 1] = 16384
    mid = oracle_run(_base(prop, s), NT)

    def absorbed(r):
        return 1 - r["genergy"][0::2].sum() / r["genergy"][1::2].sum()
    lo, hi = sorted([absorbed(ref), absorbed(ref2)])
    assert lo < absorbed(mid) < hi


def test_word_media_drop_per_medium_record_columns():
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37]]
    v = _labels()
    m = np.where(v == 1, np.float32(0.01), np.float32(np.nan))[None].astype(np.float32)
    cfg = _base(prop, m, detpos=[[10.5, 10.5, 0.0, 3.0]], issavedet=1, savedetflag="DSPX")
    ora = oracle_run(cfg, NT)
    assert ora["reclen"] == 1 + 3       # detector id + exit position


# ---- split voxels (svmc, 97) -------------------------------------------------

def _svmc_raw(lower, upper=None, point=None, normal=None):
    """int8 [8, nx, ny, nz] split-voxel input from per-voxel labels, interface
    point bytes and normal bytes (src/pmcx.cpp:125-136 layout)."""
    lower = np.asarray(lower)
    raw = np.zeros((8,) + lower.shape, dtype=np.uint8)
    raw[0] = lower
    if upper is not None:
        raw[1] = upper
    if point is not None:
        raw[2:5] = point
    if normal is not None:
        raw[5:8] = normal
    return raw.view(np.int8)


def _product_words(cfg):
    """The volume words the library hands the kernel after validation."""
    c = _lib.Config()
    from mcx_amd import pmcx
    pmcx._parse_config(cfg, c)
    L = _lib.lib()
    L.mcx_validatecfg(ctypes.byref(c), None, None, 0)
    code, msg = _lib.last_error()
    assert code == 0, msg
    n = c.dim.x * c.dim.y * c.dim.z
    w = np.ctypeslib.as_array(c.vol, shape=(2 * n,)).copy()
    L.mcx_clearcfg(ctypes.byref(c))
    return w


def test_svmc_planes_match_library_preprocessing():
    rng = np.random.default_rng(11)
    raw = _svmc_raw(rng.integers(0, 3, (6, 5, 4)), rng.integers(0, 3, (6, 5, 4)),
                    rng.integers(0, 256, (3, 6, 5, 4)), rng.integers(0, 256, (3, 6, 5, 4)))
    prop = [[0, 0, 1, 1], [0.01, 1, 0.9, 1.37], [0.02, 2, 0.5, 1.4]]
    cfg = _base(prop, raw)
    words, mb, dims = media.encode_volume(raw)
    assert mb == 97 and dims == (6, 5, 4)
    planes = media.svmc_planes(words)
    assert np.array_equal(_product_words(cfg), planes)
    # voxel (1, 2, 3): plane 1 = {py, px, upper, lower}, plane 2 = {nz, ny, nx, pz} from the low byte
    b = raw.view(np.uint8)[:, 1, 2, 3].astype(np.uint32)
    i = 1 + 6 * 2 + 30 * 3
    assert planes[i] == b[3] | b[2] << 8 | b[1] << 16 | b[0] << 24
    assert planes[120 + i] == b[7] | b[6] << 8 | b[5] << 16 | b[4] << 24
    # the detector mask: product and oracle agree
    cfg2 = _base(prop, raw, detpos=[[2.0, 2.0, 0.0, 2.0]], issavedet=1)
    import oracle_py
    want = oracle_py.maskdet(planes[:120].reshape(dims, order="F"), np.array([[2, 2, 0, 2]], np.float32), svmc=1)
    got = _product_words(cfg2)[:120]
    assert np.array_equal(got, want.ravel(order="F"))
    assert np.any(got >> 31)


def _two_labels():
    v = np.ones((21, 21, 21), dtype=np.uint8)
    v[:, :, 11:] = 2
    v[:2] = 0
    return v


def test_svmc_unsplit_voxels_equal_label_volume():
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.03, 2.0, 0.5, 1.5]]
    v = _two_labels()
    ref = oracle_run(_base(prop, v), NT)
    _same(ref, oracle_run(_base(prop, _svmc_raw(v)), NT))


def test_svmc_plane_outside_the_voxel_equals_label_volume():
    """Every voxel split by a plane through its low corner with the normal
    pointing away from it: the whole voxel is on the lower side, the upper
    label is never used and the interface is never crossed."""
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37], [0.03, 2.0, 0.5, 1.5]]
    v = _two_labels()
    ref = oracle_run(_base(prop, v), NT)
    raw = _svmc_raw(v, np.where(v > 0, 3 - v, 0), np.zeros((3,) + v.shape), np.zeros((3,) + v.shape))
    _same(ref, oracle_run(_base(prop, raw), NT))
    # the same plane with the normal reversed puts the whole voxel on the upper side
    raw2 = _svmc_raw(v, np.where(v > 0, 3 - v, 0), np.zeros((3,) + v.shape), np.full((3,) + v.shape, 255))
    flipped = np.where(v > 0, 3 - v, 0).astype(np.uint8)
    ref2 = oracle_run(_base(prop, flipped), NT)
    assert abs(oracle_run(_base(prop, raw2), NT)["absorbed"] - ref2["absorbed"]) < 0.02


@pytest.mark.parametrize("n2", [1.37, 1.6])
def test_svmc_interface_matches_finer_label_grid(n2):
    """A layer interface at z = 10.5 voxels drawn as split voxels in layer
    z = 10 reproduces, within Monte Carlo error, a label volume of half-size
    voxels whose layer boundary falls on the face z = 21; with n2 != 1.37
    the interface refracts and reflects (Fresnel on the plane)."""
    prop = [[0, 0, 1, 1], [0.01, 1.0, 0.8, 1.37], [0.04, 3.0, 0.5, n2]]
    n = 20
    lab = np.ones((n, n, n), dtype=np.uint8)
    lab[:, :, 11:] = 2
    up = np.zeros_like(lab)
    up[:, :, 10] = 2
    pt = np.zeros((3, n, n, n), dtype=np.uint8)
    pt[:, :, :, 10] = 127
    nv = np.zeros((3, n, n, n), dtype=np.uint8)
    nv[:, :, :, 10] = np.array([128, 128, 255], dtype=np.uint8)[:, None, None]
    raw = _svmc_raw(lab, up, pt, nv)
    fine = np.ones((2 * n, 2 * n, 2 * n), dtype=np.uint8)
    fine[:, :, 21:] = 2
    kw = dict(nphoton=40000, tend=2e-8, tstep=2e-8, outputtype="energy")
    a = oracle_run(_base(prop, raw, srcpos=[10.0, 10.0, 0.0], **kw), 512)
    b = oracle_run(_base(prop, fine, srcpos=[20.0, 20.0, 0.0], unitinmm=0.5, **kw), 512)
    assert abs(a["absorbed"] - b["absorbed"]) < 0.012, (a["absorbed"], b["absorbed"])
    # share of the absorbed energy deposited below the interface
    fa = (a["field"][:, :, 11:].sum()) / a["field"].sum()
    fb = (b["field"][:, :, 22:].sum()) / b["field"].sum()
    # (a plane at z = 10.0 or 10.98 instead moves this share by 0.045)
    assert abs(fa - fb) < 0.015, (fa, fb)


@pytest.mark.parametrize("fmt,raw_of", [
    ("mua_float", lambda g: g.uniform(0.001, 0.05, (1, 6, 5, 4)).astype(np.float32)),
    ("muamus_float", lambda g: g.uniform(0.001, 3.0, (2, 6, 5, 4)).astype(np.float32)),
    ("asgn_float", lambda g: g.uniform(0.001, 1.4, (4, 6, 5, 4)).astype(np.float32)),
    ("asgn_byte", lambda g: g.integers(0, 128, (4, 6, 5, 4)).astype(np.uint8)),
    ("muamus_short", lambda g: g.integers(0, 65536, (2, 6, 5, 4)).astype(np.uint16)),
    ("svmc", lambda g: g.integers(0, 3, (8, 6, 5, 4)).astype(np.uint8)),
])
def test_volume_file_media_formats(tmp_path, fmt, raw_of):
    """Domain.MediaFormat + VolumeFile (src/mcx_utils.c:3510-3600): the
    library's file loader and preprocessing give the same words as the pmcx
    array encoding of the same data."""
    g = np.random.default_rng(3)
    arr = raw_of(g)
    if fmt == "svmc":
        arr[1] = np.where(arr[1] == arr[0], 0, arr[1])
    # the file holds the components of each voxel together, x fastest
    np.ascontiguousarray(np.asfortranarray(arr).reshape(arr.shape[0], -1, order="F").T).tofile(tmp_path / "vol.bin")
    doc = {"Session": {"ID": "m", "Photons": 1000},
           "Domain": {"Dim": [6, 5, 4], "MediaFormat": fmt, "VolumeFile": str(tmp_path / "vol.bin"), "LengthUnit": 0.5,
                      "Media": [[0, 0, 1, 1], [0.01, 1, 0.9, 1.37], [0.02, 2, 0.5, 1.37]]},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Optode": {"Source": {"Pos": [2, 2, 0], "Dir": [0, 0, 1]}}}
    c = _lib.Config()
    L = _lib.lib()
    L.mcx_initcfg(ctypes.byref(c))
    c.parentid = 2
    import json
    L.mcx_readconfig(json.dumps(doc).encode(), ctypes.byref(c))
    code, msg = _lib.last_error()
    assert code == 0, msg
    n = 120
    two = fmt in ("asgn_float", "svmc")
    got = np.ctypeslib.as_array(c.vol, shape=((2 if two else 1) * n,)).copy()
    L.mcx_clearcfg(ctypes.byref(c))
    src = arr.view(np.int8) if arr.dtype == np.uint8 else (arr.view(np.int16) if arr.dtype == np.uint16 else arr)
    want, mb, _ = media.encode_volume(src, unitinmm=0.5)
    if fmt == "svmc":
        want = media.svmc_planes(want)
    assert np.array_equal(got, want)
