"""ctypes front-end of the CPU oracle (test infrastructure only).

Builds the preprocessed inputs the reference kernel sees and runs the
oracle's restatement of the transport loop.  Only tests/, smoke() and
bench.py's cpu_baseline leg import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

DET_MASK = 0x80000000
EPS = np.float32(1.1920928955078125e-07)
R_C0 = np.float32(3.335640951981520e-12)

_f4 = ctypes.c_float * 4


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint), ("ny", ctypes.c_uint), ("nz", ctypes.c_uint),
        ("media", ctypes.POINTER(ctypes.c_uint)),
        ("prop", ctypes.POINTER(ctypes.c_float)),
        ("medianum", ctypes.c_uint),
        ("detpos", ctypes.POINTER(ctypes.c_float)),
        ("detnum", ctypes.c_uint),
        ("srctype", ctypes.c_int),
        ("srcpos", _f4), ("srcdir", _f4), ("srcparam1", _f4), ("srcparam2", _f4),
        ("twin0", ctypes.c_float), ("twin1", ctypes.c_float), ("Rtstep", ctypes.c_float),
        ("oneoverc0", ctypes.c_float), ("minaccumtime", ctypes.c_float),
        ("maxgate", ctypes.c_uint),
        ("minenergy", ctypes.c_float),
        ("isreflect", ctypes.c_int), ("doreflect", ctypes.c_int),
        ("isspecular", ctypes.c_int), ("voidtime", ctypes.c_int), ("maxvoidstep", ctypes.c_int),
        ("gscatter", ctypes.c_uint),
        ("outputtype", ctypes.c_int),
        ("save2pt", ctypes.c_int),
        ("bc", ctypes.c_ubyte * 12),
        ("is2d", ctypes.c_int),
        ("savedet", ctypes.c_int),
        ("savedetflag", ctypes.c_uint), ("maxdetphoton", ctypes.c_uint),
        ("nthread", ctypes.c_uint),
        ("threadphoton", ctypes.c_int), ("oddphotons", ctypes.c_int),
        ("seeds", ctypes.POINTER(ctypes.c_uint)),
        ("thread_begin", ctypes.c_uint), ("thread_end", ctypes.c_uint),
        ("ompthreads", ctypes.c_int),
        ("saveseed", ctypes.c_int),
        ("replayseed", ctypes.POINTER(ctypes.c_uint)),
        ("replayweight", ctypes.POINTER(ctypes.c_float)),
        ("replaytof", ctypes.POINTER(ctypes.c_float)),
        ("replaydetid", ctypes.POINTER(ctypes.c_int)),
        ("replaydet", ctypes.c_int),
        ("srcpattern", ctypes.POINTER(ctypes.c_float)),
        ("issaveref", ctypes.c_int),
        ("mediaformat", ctypes.c_int),
        ("omega", ctypes.c_float),
        ("replayindex", ctypes.c_int),
    ]


class OracleOut(ctypes.Structure):
    _fields_ = [
        ("field", ctypes.POINTER(ctypes.c_double)),
        ("field32", ctypes.POINTER(ctypes.c_float)),
        ("genergy", ctypes.POINTER(ctypes.c_float)),
        ("detdata", ctypes.POINTER(ctypes.c_float)),
        ("detected", ctypes.c_uint),
        ("reclen", ctypes.c_uint),
        ("errors", ctypes.c_int),
        ("nstep", ctypes.c_ulonglong), ("ndeposit", ctypes.c_ulonglong),
        ("nscatter", ctypes.c_ulonglong), ("nlaunch", ctypes.c_ulonglong),
        ("seconds", ctypes.c_double),
        ("seeddata", ctypes.POINTER(ctypes.c_uint)),
        ("thread_steps", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.mcx_oracle_run.argtypes = [ctypes.POINTER(OracleParams), ctypes.POINTER(OracleOut)]
        L.mcx_oracle_run.restype = ctypes.c_int
        L.mcx_oracle_glibc_rand.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_int)]
        L.mcx_oracle_seedwords.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint)]
        L.mcx_oracle_rng_floats.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_uint, ctypes.POINTER(ctypes.c_float)]
        L.mcx_oracle_maskdet.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                         ctypes.POINTER(ctypes.c_float), ctypes.c_uint, ctypes.c_int]
        L.mcx_oracle_expf.argtypes = [ctypes.c_float]
        L.mcx_oracle_expf.restype = ctypes.c_float
        L.mcx_oracle_logf.argtypes = [ctypes.c_float]
        L.mcx_oracle_logf.restype = ctypes.c_float
        L.mcx_oracle_sincos2pi.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.mcx_oracle_sincosf.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        for name in ("mcx_oracle_hg_costheta",):
            getattr(L, name).argtypes = [ctypes.c_float, ctypes.c_float]
            getattr(L, name).restype = ctypes.c_float
        for name in ("mcx_oracle_fresnel_face", "mcx_oracle_reflectcoeff"):
            getattr(L, name).argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
            getattr(L, name).restype = ctypes.c_float
        L.mcx_oracle_transmit.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float)]
        L.mcx_oracle_trace.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_uint,
                                       ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_float)]
        L.mcx_oracle_trace.restype = ctypes.c_int
        L.mcx_oracle_launch_samples.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_uint, ctypes.POINTER(ctypes.c_float)]
        L.mcx_oracle_launch_samples.restype = ctypes.c_int
        L.mcx_oracle_set_mutation.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def seedwords(seed, nwords):
    w = np.zeros(nwords, dtype=np.uint32)
    lib().mcx_oracle_seedwords(seed, nwords, _ptr(w, ctypes.c_uint))
    return w


def glibc_rand(seed, n):
    o = np.zeros(n, dtype=np.int32)
    lib().mcx_oracle_glibc_rand(seed, n, _ptr(o, ctypes.c_int))
    return o


def rng_floats(seed4, n):
    s = np.ascontiguousarray(seed4, dtype=np.uint32)
    o = np.zeros(n, dtype=np.float32)
    lib().mcx_oracle_rng_floats(_ptr(s, ctypes.c_uint), n, _ptr(o, ctypes.c_float))
    return o


def maskdet(vol, detpos, svmc=0):
    """vol: uint32 array indexed [x,y,z] (F-order); detpos: (n,4) 0-based;
    svmc: vol is the first plane of split-voxel words."""
    v = np.asfortranarray(vol, dtype=np.uint32).copy(order="F")
    flat = v.ravel(order="F")
    d = np.ascontiguousarray(detpos, dtype=np.float32)
    nx, ny, nz = v.shape
    lib().mcx_oracle_maskdet(_ptr(flat, ctypes.c_uint), nx, ny, nz, _ptr(d, ctypes.c_float), d.shape[0], int(svmc))
    return flat.reshape((nx, ny, nz), order="F")


class mutation:
    """Context manager that plants mistake `mid` in the oracle (ids in
    oracle/mcx_oracle.h) and removes it on exit."""

    def __init__(self, mid):
        self.mid = int(mid)

    def __enter__(self):
        lib().mcx_oracle_set_mutation(self.mid)
        return self

    def __exit__(self, *exc):
        lib().mcx_oracle_set_mutation(0)
        return False


def hg_costheta(g, u):
    L = lib()
    return np.array([L.mcx_oracle_hg_costheta(float(g), float(x)) for x in np.asarray(u, dtype=np.float32)],
                    dtype=np.float32)


def fresnel_face(n1, n2, cosi):
    return lib().mcx_oracle_fresnel_face(float(n1), float(n2), float(cosi))


def reflectcoeff(n1, n2, cosi):
    return lib().mcx_oracle_reflectcoeff(float(n1), float(n2), float(cosi))


def transmit(n1, n2, face, v):
    vin = np.ascontiguousarray(v, dtype=np.float32)
    vout = np.zeros(3, dtype=np.float32)
    lib().mcx_oracle_transmit(float(n1), float(n2), int(face), _ptr(vin, ctypes.c_float), _ptr(vout, ctypes.c_float))
    return vout


def trace(p0, v0, dims, maxn=4096):
    """Voxels entered and step lengths of a straight ray (the loop's stepping)."""
    p = np.ascontiguousarray(p0, dtype=np.float32)
    v = np.ascontiguousarray(v0, dtype=np.float32)
    ids = np.zeros(3 * maxn, dtype=np.int32)
    lens = np.zeros(maxn, dtype=np.float32)
    n = lib().mcx_oracle_trace(_ptr(p, ctypes.c_float), _ptr(v, ctypes.c_float), dims[0], dims[1], dims[2], maxn,
                               _ptr(ids, ctypes.c_int), _ptr(lens, ctypes.c_float))
    return ids[: 3 * n].reshape(n, 3), lens[:n]


SRCTYPES = ["pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier", "arcsine", "disk", "fourierx",
            "fourierx2d", "zgaussian", "line", "slit", "pencilarray", "pattern3d", "hyperboloid", "ring"]


def launch_samples(srctype, srcpos, srcdir, n, srcparam1=(0, 0, 0, 0), srcparam2=(0, 0, 0, 0), seed=1234,
                   srcpattern=None, dims=(64, 64, 64)):
    """n launches of one virtual thread in a uniform label-1 volume (so no
    void skipping moves them): returns positions (n,3), weights (n,),
    directions (n,3).  srcdir is normalised as the host does; its 4th
    element is the focal length."""
    st = SRCTYPES.index(srctype) if isinstance(srctype, str) else int(srctype)
    nx, ny, nz = dims
    media = np.ones(nx * ny * nz, dtype=np.uint32)
    prop = np.array([[0, 0, 1, 1], [0.01, 1.0, 0.9, 1.37]], dtype=np.float32)
    dp = np.zeros((1, 4), dtype=np.float32)
    sd = np.array(srcdir, dtype=np.float64)
    nrm = 1.0 / np.sqrt((sd[:3] ** 2).sum())
    srcdir4 = [sd[0] * np.float32(nrm), sd[1] * np.float32(nrm), sd[2] * np.float32(nrm),
               srcdir[3] if len(srcdir) > 3 else 0.0]
    sp = list(srcpos)[:3] + [srcpos[3] if len(srcpos) > 3 else 1.0]
    p2 = np.array(srcparam2, dtype=np.float32).copy()
    if st in (0, 1, 2, 7, 11):
        i1 = int(np.floor(sp[2])) * nx * ny + int(np.floor(sp[1])) * nx + int(np.floor(sp[0]))
        p2[2:4] = np.array([i1, 1], dtype=np.uint32).view(np.float32)
    seeds = seedwords(seed, 4)
    P = OracleParams()
    P.nx, P.ny, P.nz = nx, ny, nz
    P.media = _ptr(media, ctypes.c_uint)
    P.prop = _ptr(prop, ctypes.c_float)
    P.medianum = 2
    P.detpos = _ptr(dp, ctypes.c_float)
    P.detnum = 0
    P.srctype = st
    P.srcpos[:] = list(np.array(sp, dtype=np.float32))
    P.srcdir[:] = list(np.array(srcdir4, dtype=np.float32))
    P.srcparam1[:] = list(np.array(srcparam1, dtype=np.float32))
    P.srcparam2[:] = list(p2)
    P.twin0, P.twin1, P.Rtstep = 0.0, 5e-9, 2e8
    P.oneoverc0 = P.minaccumtime = R_C0
    P.maxgate = 1
    P.maxvoidstep = 1000
    P.gscatter = 1000000000
    P.save2pt = 1
    P.nthread = 1
    P.threadphoton = n
    P.seeds = _ptr(seeds, ctypes.c_uint)
    keep = [media, prop, dp, seeds]
    if srcpattern is not None:
        pat = np.ascontiguousarray(np.asarray(srcpattern, dtype=np.float32).ravel(order="F"))
        keep.append(pat)
        P.srcpattern = _ptr(pat, ctypes.c_float)
    out = np.zeros(8 * n, dtype=np.float32)
    rc = lib().mcx_oracle_launch_samples(ctypes.byref(P), n, _ptr(out, ctypes.c_float))
    if rc != n:
        raise RuntimeError("launch sampling failed (%d)" % rc)
    o = out.reshape(n, 8)
    return o[:, 0:3].astype(np.float64), o[:, 3].astype(np.float64), o[:, 4:7].astype(np.float64)


def savedet_reclen(savedetflag, medianum):
    f = savedetflag
    partial = (medianum - 1) * (((f >> 1) & 1) + ((f >> 2) & 1) + ((f >> 3) & 1))
    return partial + (f & 1) + 3 * (((f >> 4) & 1) + ((f >> 5) & 1)) + ((f >> 6) & 1)


def run(vol, prop, srcpos, srcdir, nphoton, nthread, seed, tstart=0.0, tend=5e-9, tstep=5e-9,
        srctype=0, srcparam1=(0, 0, 0, 0), srcparam2=(0, 0, 0, 0), detpos=None, isreflect=1,
        isspecular=0, unitinmm=1.0, outputtype=0, savedetflag=0x5, issavedet=None,
        maxdetphoton=1000000, minenergy=0.0, bc="", gscatter=1000000000, field32=False,
        thread_range=None, ompthreads=0, maxgate=None, seeds=None, voidtime=1, want_field=True,
        saveseed=0, replay=None, replaydet=0, srcpattern=None, issaveref=0, mediaformat=0, words2=None,
        thread_steps=False, genergy=None, omega=0.0, replayindex=0):
    """Run the oracle on a preprocessed problem.

    vol: uint8/uint32 labels indexed [x,y,z]; prop: (M,4) mua/mus/g/n per mm;
    srcpos/detpos: 0-based grid coordinates; srcdir normalised here.
    Returns a dict with field (FP64, [x,y,z,gate]), genergy, detp, counters.
    saveseed: also return the launch RNG state of each detected photon
    ("seeds", uint8 [n, 16]).  replay: dict with "seed" (uint8 [n, 16]),
    "weight", "tof", "detid" per replayed photon (n = nphoton); with
    replaydet == -1 the field has one block of gates per detector; the RF
    outputs (6 otRF, 8 otRFmus) add an imaginary plane of as many gates
    after the real one (omega: modulation angular frequency, rad/s).
    replayindex 1 numbers replayed photons as the reference does.
    mediaformat > 4: vol holds the 32-bit property words of that format
    (words2: the second plane of the two-word formats 96 and 97).
    """
    vol = np.asarray(vol)
    nx, ny, nz = vol.shape
    media = np.asfortranarray(vol.astype(np.uint32)).ravel(order="F").copy()
    prop = np.array(prop, dtype=np.float32).copy()
    for i in range(1, prop.shape[0]):
        if unitinmm != 1.0:
            prop[i, 0] *= np.float32(unitinmm)
            prop[i, 1] *= np.float32(unitinmm)
    for i in range(prop.shape[0]):
        if prop[i, 1] == 0.0:
            prop[i, 1] = EPS
            prop[i, 2] = 1.0
    detnum = 0 if detpos is None else len(detpos)
    if issavedet is None:
        issavedet = 1 if detnum > 0 else 0
    if not issavedet:
        savedetflag = 0
    dp = np.zeros((max(detnum, 1), 4), dtype=np.float32)
    if detnum:
        dp[:detnum] = np.array(detpos, dtype=np.float32)
        if issavedet:
            media = maskdet(media.reshape((nx, ny, nz), order="F"), dp[:detnum],
                            mediaformat == 97).ravel(order="F").copy()
    if mediaformat > 4 and mediaformat != 97:
        # property-word media keep no per-medium record columns
        savedetflag &= ~(2 | 4 | 8 | 128)
    if mediaformat in (96, 97):
        media = np.concatenate([media, np.asfortranarray(np.asarray(words2, dtype=np.uint32)).ravel(order="F")])
    sd = np.array(srcdir, dtype=np.float64)
    nrm = 1.0 / np.sqrt((sd[:3] ** 2).sum())
    srcdir4 = np.array([sd[0] * np.float32(nrm), sd[1] * np.float32(nrm), sd[2] * np.float32(nrm),
                        srcdir[3] if len(srcdir) > 3 else 0.0], dtype=np.float32)
    sp = np.array(list(srcpos)[:3] + [srcpos[3] if len(srcpos) > 3 else 1.0], dtype=np.float32)
    if sp[3] == 0:
        sp[3] = 1.0
    p2 = np.array(srcparam2, dtype=np.float32)
    if srctype in (0, 1, 2, 7, 11):
        if (sp[:3] < 0).any() or sp[0] >= nx or sp[1] >= ny or sp[2] >= nz:
            i1, m1 = 0, 0
        else:
            i1 = int(np.floor(sp[2])) * nx * ny + int(np.floor(sp[1])) * nx + int(np.floor(sp[0]))
            m1 = int(media[i1] & 0x7FFFFFFF)
        p2 = p2.copy()
        p2[2:4] = np.array([i1, m1], dtype=np.uint32).view(np.float32)
    gates = int((tend - tstart) / tstep + 0.5)
    if maxgate is None:
        maxgate = gates
    if seeds is None:
        seeds = seedwords(seed, nthread * 4)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    threadphoton = nphoton // nthread
    oddphotons = nphoton - threadphoton * nthread
    P = OracleParams()
    P.nx, P.ny, P.nz = nx, ny, nz
    P.media = _ptr(media, ctypes.c_uint)
    P.prop = _ptr(prop, ctypes.c_float)
    P.medianum = prop.shape[0]
    P.detpos = _ptr(dp, ctypes.c_float)
    P.detnum = detnum
    P.srctype = srctype
    P.srcpos[:] = list(sp)
    P.srcdir[:] = list(srcdir4)
    P.srcparam1[:] = list(np.array(srcparam1, dtype=np.float32))
    P.srcparam2[:] = list(p2)
    P.twin0 = tstart
    P.twin1 = np.float32(tstart) + np.float32(tstep) * np.float32(maxgate)
    P.Rtstep = np.float32(1.0) / np.float32(tstep)
    P.oneoverc0 = R_C0 * np.float32(unitinmm)
    P.minaccumtime = np.float32(1.0) * R_C0 * np.float32(unitinmm)
    P.maxgate = maxgate
    P.minenergy = minenergy
    bcv = np.zeros(12, dtype=np.uint8)
    code = {"_": 0, "r": 1, "a": 2, "m": 3, "c": 4, "0": 0, "1": 1}
    for i, ch in enumerate(bc[:12]):
        bcv[i] = code.get(ch, 0)
    isref = isreflect or any(bcv[i] in (1, 3) for i in range(6))
    P.isreflect = int(bool(isref))
    P.doreflect = int(isreflect)
    P.isspecular = isspecular
    P.voidtime = voidtime
    P.maxvoidstep = 1000
    P.gscatter = gscatter
    P.outputtype = outputtype
    P.save2pt = 1
    P.bc[:] = list(bcv)
    # a domain with exactly one axis of extent 1 is simulated in 2-D
    # (src/mcx_core.cu:2850, :2874: is2d = the singular axis, 1-based)
    is2d = 1 if nx == 1 else (2 if ny == 1 else (3 if nz == 1 else 0))
    P.is2d = is2d if (nx > 1) + (ny > 1) + (nz > 1) == 2 else 0
    P.savedet = issavedet
    P.savedetflag = savedetflag
    P.maxdetphoton = maxdetphoton
    P.nthread = nthread
    P.threadphoton = threadphoton
    P.oddphotons = oddphotons
    P.seeds = _ptr(seeds, ctypes.c_uint)
    if thread_range is None:
        thread_range = (0, nthread)
    P.thread_begin, P.thread_end = thread_range
    P.ompthreads = ompthreads
    P.saveseed = int(saveseed)
    P.issaveref = int(issaveref)
    P.mediaformat = int(mediaformat)
    P.omega = float(omega)
    P.replayindex = int(replayindex)
    keep = []
    if srcpattern is not None:
        pat = np.ascontiguousarray(np.asarray(srcpattern, dtype=np.float32).ravel(order="F"))
        keep.append(pat)
        P.srcpattern = _ptr(pat, ctypes.c_float)
    blocks = 1
    if replay is not None:
        rs = np.ascontiguousarray(np.asarray(replay["seed"], dtype=np.uint8).reshape(-1, 16)).view(np.uint32).copy()
        rw = np.ascontiguousarray(replay["weight"], dtype=np.float32)
        rt = np.ascontiguousarray(replay["tof"], dtype=np.float32)
        rd = np.ascontiguousarray(replay["detid"], dtype=np.int32)
        keep += [rs, rw, rt, rd]
        P.replayseed = _ptr(rs, ctypes.c_uint)
        P.replayweight = _ptr(rw, ctypes.c_float)
        P.replaytof = _ptr(rt, ctypes.c_float)
        P.replaydetid = _ptr(rd, ctypes.c_int)
        P.replaydet = int(replaydet)
        if replaydet == -1:
            blocks = max(detnum, 1)
    if replay is not None and outputtype in (6, 8):
        blocks *= 2
    nvox = nx * ny * nz
    field = np.zeros(nvox * maxgate * blocks, dtype=np.float64) if want_field else None
    f32 = np.zeros(2 * nvox * maxgate * blocks, dtype=np.float32) if field32 else None
    # per-thread escaped/launched energies continue from `genergy` (respins)
    genergy = np.zeros(2 * nthread, dtype=np.float32) if genergy is None else np.array(genergy, dtype=np.float32)
    reclen = savedet_reclen(savedetflag, prop.shape[0])
    detdata = np.zeros(max(maxdetphoton, 1) * max(reclen, 1), dtype=np.float32) if issavedet else None
    O = OracleOut()
    O.field = _ptr(field, ctypes.c_double) if field is not None else None
    O.field32 = _ptr(f32, ctypes.c_float) if f32 is not None else None
    O.genergy = _ptr(genergy, ctypes.c_float)
    O.detdata = _ptr(detdata, ctypes.c_float) if detdata is not None else None
    seedbuf = np.zeros(max(maxdetphoton, 1) * 4, dtype=np.uint32) if (saveseed and issavedet) else None
    O.seeddata = _ptr(seedbuf, ctypes.c_uint) if seedbuf is not None else None
    tsteps = np.zeros(nthread, dtype=np.uint64) if thread_steps else None
    O.thread_steps = _ptr(tsteps, ctypes.c_ulonglong) if tsteps is not None else None
    rc = lib().mcx_oracle_run(ctypes.byref(P), ctypes.byref(O))
    if rc != 0:
        raise RuntimeError("oracle rejected input (%d)" % rc)
    out = {
        "genergy": genergy,
        "energyesc": float(genergy[0::2].astype(np.float64).sum()),
        "energytot": float(genergy[1::2].astype(np.float64).sum()),
        "detected": int(O.detected),
        "reclen": int(O.reclen),
        "errors": int(O.errors),
        "nstep": int(O.nstep), "ndeposit": int(O.ndeposit),
        "nscatter": int(O.nscatter), "nlaunch": int(O.nlaunch),
        "seconds": float(O.seconds),
        "media": media,
        "maxgate": maxgate,
    }
    out["absorbed"] = (out["energytot"] - out["energyesc"]) / out["energytot"] if out["energytot"] > 0 else 0.0
    if field is not None:
        out["field"] = field.reshape((nx, ny, nz, maxgate * blocks), order="F")
    if f32 is not None:
        raw = f32[: nvox * maxgate * blocks] + f32[nvox * maxgate * blocks:]
        out["field32"] = raw.reshape((nx, ny, nz, maxgate * blocks), order="F")
    if tsteps is not None:
        out["thread_steps"] = tsteps
    if seedbuf is not None:
        n = min(out["detected"], maxdetphoton)
        out["seeds"] = seedbuf[: 4 * n].view(np.uint8).reshape(n, 16).copy()
    if detdata is not None:
        n = min(out["detected"], maxdetphoton)
        out["detp"] = detdata[: n * reclen].reshape((n, reclen)) if reclen else np.zeros((n, 0), np.float32)
    return out
