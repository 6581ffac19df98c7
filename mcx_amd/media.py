"""Volume encodings of the pmcx `vol` array (src/pmcx.cpp:102-380).

A 2-D/3-D integer array is a label volume.  A 4-D array whose first axis
holds per-voxel components selects a property-word format by its dtype and
component count:

    int8/uint8   [4, ...]  asgn_byte    (103) bytes mua, mus, g, n fractions
    int8/uint8   [8, ...]  svmc         (97)  split voxels: lower, upper label,
                                              interface point x, y, z, normal x, y, z
    int16/uint16 [3, ...]  mixlabel     (98)  label0, label1, label1 fraction
    int16/uint16 [2, ...]  muamus_short (104) shorts mua, mus fractions
    float        [1, ...]  mua_float    (101) mua
    float        [2, ...]  muamus_float (100) mua, mus -> half words
    float        [3, ...]  labelplus    (99)  value, property index, label
    float        [4, ...]  asgn_float   (96)  mua, mus, g, n -> 2 half words

`encode_volume` returns the 32-bit words in x-fastest order (two planes for
asgn_float; the raw 8 bytes per voxel for svmc, which the library's
preprocessing turns into two planes, see `svmc_planes`), the mediabyte
value and the grid size.  The half-precision
encoders reproduce the reference's bit rules (mcx_float2half2,
src/mcx_utils.c:5490-5546, and the inline encoder of src/pmcx.cpp:291-295),
since those rules define the stored properties.
"""
import numpy as np

MEDIA_ASGN_F2H = 96
MEDIA_2LABEL_SPLIT = 97
MEDIA_2LABEL_MIX = 98
MEDIA_LABEL_HALF = 99
MEDIA_AS_F2H = 100
MEDIA_MUA_FLOAT = 101
MEDIA_AS_HALF = 102
MEDIA_ASGN_BYTE = 103
MEDIA_AS_SHORT = 104

FLT_EPSILON = np.float32(1.1920928955078125e-07)


def _bits(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)


def _half_exp(b):
    e = (b >> 23) & 0xFF
    return np.where(e > 0x70, (e - 0x70) & 0x1F, 0)


def _half_denorm(b):
    e = (b >> 23) & 0xFF
    m = ((b >> 12) & 0x7FF) | 0x800
    s1 = 114 - e.astype(np.int64)
    s2 = 113 - e.astype(np.int64)
    q1 = np.where(s1 < 32, m >> np.minimum(s1, 63).astype(np.uint64), 0)
    q2 = np.where(s2 < 32, (m >> np.minimum(s2, 63).astype(np.uint64)) & 1, 0)
    return ((b >> 16) & 0x8000) | ((q1 + q2) & 0xFFFF)


def float2half2(a, b):
    """Packed half words of (a, b) by the reference encoder's rules: the
    mantissa is truncated; small exponents take a rounded denormal path
    only when the top mantissa bits are also small; the three lowest
    mantissa bits of `a` come from the sign/exponent bits already written
    (the encoder's union aliasing); a word of 0 becomes 0x00010000 and a
    word of 0x80000000 becomes 0."""
    ba, bb = _bits(a), _bits(b)
    ta, tb = _half_exp(ba), _half_exp(bb)
    top = ((((ba >> 31) << 5) | ta) << 10) & 0xFFFF
    h0 = np.where((((ba >> 13) & 0x3FF) < 0x10) & (ta == 0), _half_denorm(ba),
                  top | ((ba >> 13) & 0x3F8) | ((top >> 13) & 7))
    h1 = np.where((((bb >> 13) & 0x3FF) < 0x10) & (tb == 0), _half_denorm(bb),
                  (((((bb >> 31) << 5) | tb) << 10) & 0xFFFF) | ((bb >> 13) & 0x3FF))
    w = (h0 & 0xFFFF) | ((h1 & 0xFFFF) << 16)
    w = np.where(w == 0, 0x00010000, w)
    w = np.where(w == 0x80000000, 0, w)
    return w.astype(np.uint32)


def half_truncate(x):
    """src/pmcx.cpp:291-295: sign, rebased exponent (0 at or below 2^-15)
    and the top 10 mantissa bits of a float."""
    b = _bits(x)
    return (((((b >> 31) << 5) | _half_exp(b)) << 10) & 0xFFFF) | ((b >> 13) & 0x3FF)


def _components(vol, n):
    """[n, nx, ny, nz] array -> (n, nvox) in x-fastest voxel order."""
    return np.asfortranarray(vol).reshape(n, -1, order="F")


def encode_volume(vol, unitinmm=1.0):
    """pmcx `vol` -> (words uint32, mediabyte, (nx, ny, nz))."""
    vol = np.asarray(vol)
    if vol.ndim == 2:
        vol = vol[:, :, None]
    u = np.float32(unitinmm)
    if vol.ndim == 3:
        if vol.dtype in (np.int8, np.uint8, np.bool_):
            mb = 1
        elif vol.dtype in (np.int16, np.uint16):
            mb = 2
        elif vol.dtype in (np.int32, np.uint32, np.int64, np.uint64):
            mb = 4
        elif vol.dtype in (np.float32, np.float64):
            if not np.all(vol == np.round(vol)):
                raise ValueError("a 3-D floating-point volume must hold integer labels; "
                                 "give per-voxel properties as a 4-D array")
            mb = 4
        else:
            raise ValueError("unsupported volume data type %s" % vol.dtype)
        words = np.asfortranarray(vol).ravel(order="F").astype(np.uint32)
        return words, mb, tuple(int(s) for s in vol.shape)
    if vol.ndim != 4:
        raise ValueError("Configuration must specify a 2/3/4D volume.")
    nc = vol.shape[0]
    dims = tuple(int(s) for s in vol.shape[1:])
    if vol.dtype in (np.int8, np.uint8):
        if nc == 4:
            c = _components(vol.astype(np.uint8), 4).astype(np.uint32)
            return (c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24)).astype(np.uint32), MEDIA_ASGN_BYTE, dims
        if nc == 8:
            # raw 8 bytes per voxel; the library turns them into two planes
            raw = np.ascontiguousarray(_components(vol.astype(np.uint8), 8).T)
            return raw.view(np.uint32).ravel().copy(), MEDIA_2LABEL_SPLIT, dims
        raise ValueError("Invalid volume dims for int8 volume.")
    if vol.dtype in (np.int16, np.uint16):
        c = _components(vol.astype(np.uint16), nc).astype(np.uint32)
        if nc == 3:
            w = (c[0] & 0xFF) | ((c[1] & 0xFF) << 8) | ((c[2] & 0x7FFF) << 16)
            return w.astype(np.uint32), MEDIA_2LABEL_MIX, dims
        if nc == 2:
            return (c[0] | (c[1] << 16)).astype(np.uint32), MEDIA_AS_SHORT, dims
        raise ValueError("Invalid volume dims for int16 volume.")
    if vol.dtype in (np.float32, np.float64):
        c = _components(vol.astype(np.float32), nc)
        if nc == 3:
            value, index, label = c[0].copy(), c[1], c[2]
            if np.any(index < 0) or np.any(index >= 4) or np.any(label < 0):
                raise ValueError("the 2nd volume must have an integer value between 0 and 3")
            # the reference scales the value when the index is 1 or 2
            scaled = (index >= 1) & (index <= 2)
            value[scaled] = value[scaled] * u
            s0 = ((index.astype(np.uint8).astype(np.uint32) & 3) << 14) | (label.astype(np.uint16).astype(np.uint32))
            w = (s0 & 0xFFFF) | (half_truncate(value).astype(np.uint32) << 16)
            return w.astype(np.uint32), MEDIA_LABEL_HALF, dims
        if nc in (2, 4):
            mua, mus = c[0] * u, c[1] * u
            void = np.isnan(mua) | np.isnan(mus)
            w0 = np.where(void, 0, float2half2(mua, mus)).astype(np.uint32)
            if nc == 2:
                return w0, MEDIA_AS_F2H, dims
            return np.concatenate([w0, float2half2(c[2], c[3])]), MEDIA_ASGN_F2H, dims
        if nc == 1:
            f = (c[0] * u).astype(np.float32)
            w = f.view(np.uint32).copy()
            w[w == 0] = FLT_EPSILON.view(np.uint32)
            w[np.isnan(c[0])] = 0
            return w, MEDIA_MUA_FLOAT, dims
        raise ValueError("Invalid array for vol array.")
    raise ValueError("unsupported volume data type %s" % vol.dtype)


def half_words(mua, mus):
    """Raw muamus_half (102) words from values representable in binary16."""
    a = np.asarray(mua, dtype=np.float16).view(np.uint16).astype(np.uint32)
    b = np.asarray(mus, dtype=np.float16).view(np.uint16).astype(np.uint32)
    return (a | (b << 16)).astype(np.uint32)


def svmc_planes(raw_words):
    """Split-voxel words as the kernel reads them (src/mcx_utils.c:1714-1738):
    from the raw 8 bytes {lower, upper, px, py, pz, nx, ny, nz} of each
    voxel, plane 1 holds {py, px, upper, lower} and plane 2 {nz, ny, nx, pz}
    from the low byte up.  Returns both planes concatenated."""
    b = np.ascontiguousarray(raw_words, dtype=np.uint32).view(np.uint8).reshape(-1, 8).astype(np.uint32)
    p1 = b[:, 3] | (b[:, 2] << 8) | (b[:, 1] << 16) | (b[:, 0] << 24)
    p2 = b[:, 7] | (b[:, 6] << 8) | (b[:, 5] << 16) | (b[:, 4] << 24)
    return np.concatenate([p1, p2]).astype(np.uint32)
