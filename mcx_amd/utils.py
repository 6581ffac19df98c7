"""pmcx post-processing helpers for the detected-photon buffer and the
analytic diffusion checks (the reference's pmcx/pmcx/utils.py surface:
detweight :129, meanpath :97, meanscat :175, dettpsf :198, dettime :245,
tddiffusion :289, getdistance :338, detphoton :367, cwdiffusion :1044; and
the .mch reader of mcxlab/loadmch.m).

Partial path lengths in `ppath` are in grid units (multiply by unitinmm
for mm); weights are recomputed by Beer-Lambert along those paths.
"""
import re
import struct

import numpy as np

R_C0 = 3.335640951981520e-12  # 1 / speed of light in vacuum, s/mm
C0 = 299792458000.0           # speed of light in vacuum, mm/s


def _prop_of(detp, prop):
    if prop is None:
        if isinstance(detp, dict) and "prop" in detp:
            prop = detp["prop"]
        else:
            raise ValueError('Must provide input "prop"')
    prop = np.asarray(prop, dtype=np.float64)
    if prop.ndim != 2 or prop.shape[0] <= 1:
        raise ValueError("Empty property list")
    return prop


def _unit_of(detp, unitinmm):
    if unitinmm is not None:
        return float(unitinmm)
    return float(detp.get("unitinmm", 1.0)) if isinstance(detp, dict) else 1.0


def detweight(detp, prop=None, unitinmm=None):
    """Detected weight of each photon: w0 * prod_m exp(-mua_m * ppath_m * unitinmm)."""
    prop = _prop_of(detp, prop)
    if not isinstance(detp, dict) or "ppath" not in detp:
        raise ValueError('The first input must be a dict with a key named "ppath"')
    u = _unit_of(detp, unitinmm)
    ppath = np.atleast_2d(np.asarray(detp["ppath"], dtype=np.float64))
    nm = prop.shape[0] - 1
    if ppath.shape[1] < nm:
        raise ValueError("ppath has fewer columns than the property list has media")
    w0 = np.asarray(detp["w0"], dtype=np.float64).reshape(-1) if "w0" in detp else np.ones(ppath.shape[0])
    return w0 * np.exp(-(ppath[:, :nm] @ prop[1:, 0]) * u)


def dettime(detp, prop=None, unitinmm=None):
    """Time of flight of each photon, [1, n]: sum_m n_m * ppath_m * unitinmm / c0."""
    prop = _prop_of(detp, prop)
    if not isinstance(detp, dict) or "ppath" not in detp:
        raise ValueError('the first input must be a struct with a subfield named "ppath"')
    u = _unit_of(detp, unitinmm)
    ppath = np.atleast_2d(np.asarray(detp["ppath"], dtype=np.float64))
    nm = prop.shape[0] - 1
    return ((ppath[:, :nm] @ prop[1:, 3]) * R_C0 * u)[None, :]


def meanpath(detp, prop=None):
    """Weight-averaged partial path length per medium (mm)."""
    u = _unit_of(detp, None)
    w = detweight(detp, prop)
    ppath = np.atleast_2d(np.asarray(detp["ppath"], dtype=np.float64))
    return (ppath * u * w[:, None]).sum(axis=0) / w.sum()


def meanscat(detp, prop):
    """Weight-averaged total scattering count."""
    w = detweight(detp, prop)
    return float((np.asarray(detp["nscat"], dtype=np.float64) * w[:, None]).sum() / w.sum())


def dettpsf(detp, detnum, prop, time):
    """Temporal point spread function of detector `detnum`: detected weight
    summed per time bin; time = [tstart, tend, tstep], arrival times outside
    the window go to the first/last bin."""
    sel = np.asarray(detp["detid"]).reshape(-1) == detnum
    sub = {"ppath": np.atleast_2d(np.asarray(detp["ppath"]))[sel]}
    if "w0" in detp:
        sub["w0"] = np.asarray(detp["w0"]).reshape(-1)[sel]
    if "unitinmm" in detp:
        sub["unitinmm"] = detp["unitinmm"]
    w = detweight(sub, prop)
    t = dettime(sub, prop)[0]
    nbin = int(np.round((time[1] - time[0]) / time[2]))
    k = np.clip(np.ceil((t - time[0]) / time[2]), 1, nbin).astype(int) - 1
    tpsf = np.zeros((nbin, 1))
    np.add.at(tpsf[:, 0], k, w)
    return tpsf


def getdistance(srcpos, detpos):
    """separation[r, s] = |src_s - det_r|."""
    s = np.atleast_2d(np.asarray(srcpos, dtype=np.float64))
    d = np.atleast_2d(np.asarray(detpos, dtype=np.float64))
    return np.sqrt(((d[:, None, :] - s[None, :, :]) ** 2).sum(axis=2))


def _image_pair(mua, musp, Reff, srcpos):
    D = 1.0 / (3.0 * (mua + musp))
    zb = (1.0 + Reff) / (1.0 - Reff) * 2.0 * D
    z0 = 1.0 / (musp + mua)
    s = np.atleast_2d(np.asarray(srcpos, dtype=np.float64)).copy()
    real, image = s.copy(), s.copy()
    real[:, 2] += z0
    image[:, 2] -= z0 + 2.0 * zb
    return D, real, image


def tddiffusion(mua, musp, v, Reff, srcpos, detpos, t):
    """Semi-infinite time-resolved diffusion solution, 1/(mm^2 s)."""
    D, real, image = _image_pair(mua, musp, Reff, srcpos)
    r, r2 = getdistance(real, detpos), getdistance(image, detpos)
    s = 4.0 * D * v * np.asarray(t, dtype=np.float64)
    return v / (s * np.pi) ** 1.5 * np.exp(-mua * v * np.asarray(t)) * (np.exp(-r ** 2 / s) - np.exp(-r2 ** 2 / s))


def cwdiffusion(mua, musp, Reff, srcpos, detpos):
    """Semi-infinite CW diffusion solution, 1/mm^2; returns (Phi, r)."""
    D, real, image = _image_pair(mua, musp, Reff, srcpos)
    r, r2 = getdistance(real, detpos), getdistance(image, detpos)
    b = np.sqrt(3.0 * mua * musp)
    return 1.0 / (4.0 * np.pi * D) * (np.exp(-b * r) / r - np.exp(-b * r2) / r2), r


def detphoton(detp, medianum, savedetflag, issaveref=None, srcnum=None):
    """Split a raw detected-photon array [reclen, n] into named fields by
    savedetflag letters d s p m x v w i (in that record order)."""
    detp = np.atleast_2d(np.asarray(detp))
    flag = str(savedetflag)
    out = {}
    c = 0
    if re.search("[dD]", flag):
        if issaveref is not None and issaveref > 1:
            out["w0"] = detp[0, :].T
        else:
            detid = detp[0, :].astype(np.int64)
            if np.any(detid > 65535):
                out["srcid"] = (detid >> 16).astype(np.int32)
                detid = (detid & 0xFFFF).astype(np.int32)
            out["detid"] = detid.T
        c = 1
    for letter, key, n in (("sS", "nscat", medianum), ("pP", "ppath", medianum), ("mM", "mom", medianum),
                           ("xX", "p", 3), ("vV", "v", 3), ("wW", "w0", 1), ("iI", "s", 4)):
        if re.search("[%s]" % letter, flag):
            block = detp[c:c + n, :].T
            if key == "nscat":
                block = block.astype(np.int64)
            if key == "w0" and srcnum is not None and srcnum > 1:
                block = np.ascontiguousarray(block, dtype=np.float32).view(np.uint32)
            out[key] = block
            c += n
    return out


_MCH_FIELDS = ("magic", "version", "maxmedia", "detnum", "colcount", "totalphoton", "detected", "savedphoton",
               "unitinmm", "seedbyte", "normalizer", "respin", "srcnum", "savedetflag", "totalsource")


def loadmch(fname):
    """Read a .mch file: (data [savedphoton, colcount] float32, header dict,
    seeds [savedphoton, seedbyte] uint8 or None).  Concatenated files (one
    header per block) are read block by block."""
    raw = open(fname, "rb").read()
    pos, blocks, seeds, header = 0, [], [], None
    while pos + 64 <= len(raw):
        vals = struct.unpack_from("<4s7If I f i 3I i", raw, pos)
        h = dict(zip(_MCH_FIELDS, vals))
        if h["magic"] != b"MCXH":
            raise ValueError("%s is not a .mch file (bad magic at byte %d)" % (fname, pos))
        pos += 64
        n, cols = h["savedphoton"], h["colcount"]
        blocks.append(np.frombuffer(raw, dtype="<f4", count=n * cols, offset=pos).reshape(n, cols))
        pos += 4 * n * cols
        if h["seedbyte"] > 0:
            seeds.append(np.frombuffer(raw, dtype=np.uint8, count=n * h["seedbyte"], offset=pos).reshape(n, -1))
            pos += n * h["seedbyte"]
        header = h if header is None else header
    if header is None:
        raise ValueError("%s holds no .mch block" % fname)
    data = np.concatenate(blocks) if blocks else np.zeros((0, 0), np.float32)
    return data, header, (np.concatenate(seeds) if seeds else None)
