"""pmcx-compatible Python interface: run(cfg) -> dict, gpuinfo(), version().

Mirrors the reference's binding (src/pmcx.cpp): parse_config (:400-1052)
turns the dict into a Config through the C ABI, mcx_list_gpu (:1089) and
mcx_validatecfg (:1096) prepare it, mcx_run_simulation (:1143-1170) runs the
HIP engine, and the result dict carries the same keys and array layouts:
flux float32 [srcnum*Nx, Ny, Nz, Ngates] (F-order), detp float32
[reclen, ndetected] (F-order), vol (with dumpmask), and stat {runtime,
nphoton, energytot, energyabs, normalizer, unitinmm, workload}.
Errors raise ValueError for bad fields and RuntimeError for engine errors,
as the reference does (:1327-1351).
"""
import ctypes

import numpy as np

from . import _lib, media
from ._lib import Config, Float4, GPUInfo, Medium, mpPython

_libc = ctypes.CDLL(None)
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.calloc.restype = ctypes.c_void_p
_libc.calloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
_c_stderr = ctypes.c_void_p.in_dll(_libc, "stderr")

SRCTYPES = ["pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier", "arcsine", "disk", "fourierx",
            "fourierx2d", "zgaussian", "line", "slit", "pencilarray", "pattern3d", "hyperboloid", "ring"]
OUTPUTTYPES = ["flux", "fluence", "energy", "jacobian", "nscat", "wl", "wp", "wm", "rf", "length", "rfmus", "wltof",
               "wptof"]
SAVEFLAGS = "DSPMXVWI"
SEEDBYTE = 16               # bytes of one saved xorshift128+ state
MCX_SEED_FROM_FILE = -999
OT_RF, OT_RFMUS = 6, 8  # TOutputType values of the RF replay outputs
DEBUGFLAGS = "RMP"

_INT_FIELDS = ["nphoton", "nblocksize", "nthread", "maxdetphoton", "maxgate", "respin", "isreflect", "isref3",
               "isrefint", "isnormalized", "issrcfrom0", "autopilot", "printnum", "voidtime", "issaveref",
               "replaydet", "maxvoidstep", "maxjumpdebug", "gscatter", "srcnum", "srcid", "issave2pt", "issavedet",
               "issaveseed"]
_BOOL_FIELDS = ["issaveexit", "ismomentum", "isspecular", "istrajstokes", "faststep"]
_FLOAT_FIELDS = ["tstart", "tstep", "tend", "sradius", "minenergy", "unitinmm", "omega"]


def _cmalloc(nbytes):
    p = _libc.malloc(max(int(nbytes), 1))
    if not p:
        raise MemoryError("malloc failed")
    return p


def _copy_to_c(arr, ctype):
    a = np.ascontiguousarray(arr)
    p = _cmalloc(a.nbytes)
    ctypes.memmove(p, a.ctypes.data, a.nbytes)
    return ctypes.cast(p, ctypes.POINTER(ctype))


def _set_vec(dst, values, n, name):
    v = np.asarray(values, dtype=np.float64).ravel()
    if v.size < 3 or v.size > n:
        raise ValueError("the '%s' field must have 3%s elements" % (name, " or 4" if n == 4 else ""))
    for i, k in enumerate("xyzw"[: v.size]):
        setattr(dst, k, float(v[i]))


def _check():
    code, msg = _lib.last_error()
    if code:
        _lib.lib().mcx_clear_error()
        raise RuntimeError("MCX ERROR(%d): %s" % (code, msg))


def _parse_config(user_cfg, cfg):
    L = _lib.lib()
    L.mcx_initcfg(ctypes.byref(cfg))
    cfg.parentid = mpPython
    cfg.flog = None if user_cfg.get("quiet", False) else _c_stderr.value
    for k in _INT_FIELDS:
        if k in user_cfg:
            setattr(cfg, k, int(user_cfg[k]))
    for k in _BOOL_FIELDS:
        if k in user_cfg:
            setattr(cfg, k, 1 if bool(user_cfg[k]) else 0)
    for k in _FLOAT_FIELDS:
        if k in user_cfg:
            setattr(cfg, k, float(user_cfg[k]))
    if "lambda" in user_cfg:
        cfg.lambda_ = float(user_cfg["lambda"])
    if "steps" in user_cfg:
        _set_vec(cfg.steps, user_cfg["steps"], 3, "steps")
    if "vol" not in user_cfg:
        raise ValueError("Configuration must specify a 2/3/4D volume.")
    vol = np.asarray(user_cfg["vol"])
    if "mediabyte" in user_cfg and int(user_cfg["mediabyte"]) > 4:
        # raw 32-bit property words (e.g. muamus_half, 102) given as is
        mb = int(user_cfg["mediabyte"])
        if vol.ndim not in (2, 3) or mb in (media.MEDIA_ASGN_F2H, media.MEDIA_2LABEL_SPLIT):
            raise ValueError("raw media words must be a 2D/3D uint32 array of a one-word format")
        if vol.ndim == 2:
            vol = vol[:, :, None]
        flat, dims = np.asfortranarray(vol).ravel(order="F").astype(np.uint32), vol.shape
    else:
        flat, mb, dims = media.encode_volume(vol, cfg.unitinmm)
    cfg.mediabyte = mb
    cfg.dim.x, cfg.dim.y, cfg.dim.z = (int(s) for s in dims)
    cfg.vol = _copy_to_c(flat, ctypes.c_uint)
    if "srcpos" in user_cfg:
        sp = np.asarray(user_cfg["srcpos"], dtype=np.float64)
        if sp.ndim > 1:
            if sp.shape[0] > 1:
                raise ValueError("multiple sources are not supported by this engine")
            sp = sp[0]
        if sp.size < 3 or sp.size > 4:
            raise ValueError("the 'srcpos' field must have 3 or 4 columns (x,y,z,w0)")
        _set_vec(cfg.srcpos, sp, 4, "srcpos")
    if "srcdir" in user_cfg:
        sd = np.asarray(user_cfg["srcdir"], dtype=np.float64).ravel()
        if sd.size < 3 or sd.size > 4:
            raise ValueError("the 'srcdir' field must have 3 or 4 columns")
        _set_vec(cfg.srcdir, sd, 4, "srcdir")
    for k in ("srcparam1", "srcparam2", "srciquv"):
        if k in user_cfg:
            v = np.asarray(user_cfg[k], dtype=np.float64).ravel()
            dst = getattr(cfg, k)
            for i, c in enumerate("xyzw"[: min(4, v.size)]):
                setattr(dst, c, float(v[i]))
    if "detpos" in user_cfg:
        dp = np.atleast_2d(np.asarray(user_cfg["detpos"], dtype=np.float32))
        if dp.size:
            if dp.shape[1] < 3 or dp.shape[1] > 4:
                raise ValueError("the 'detpos' field must have 3 or 4 columns (x,y,z,radius)")
            d4 = np.zeros((dp.shape[0], 4), dtype=np.float32)
            d4[:, : dp.shape[1]] = dp
            if dp.shape[1] == 3:
                d4[:, 3] = cfg.detradius
            cfg.detnum = dp.shape[0]
            cfg.detpos = _copy_to_c(d4, Float4)
    if "prop" in user_cfg:
        pr = np.atleast_2d(np.asarray(user_cfg["prop"], dtype=np.float32))
        if pr.shape[1] != 4:
            raise ValueError("the 'prop' field must have 4 columns (mua,mus,g,n)")
        cfg.medianum = pr.shape[0]
        cfg.prop = _copy_to_c(pr, Medium)
    if "polprop" in user_cfg:
        raise ValueError("polarized photon simulation is not supported by this engine")
    if "session" in user_cfg:
        s = str(user_cfg["session"])
        if not s or len(s) > 255:
            raise ValueError("the 'session' field must be a non-empty string")
        cfg.session = s.encode()
    if "srctype" in user_cfg:
        s = str(user_cfg["srctype"]).lower()
        if s not in SRCTYPES:
            raise ValueError("the specified source type is not supported")
        cfg.srctype = SRCTYPES.index(s)
    if "outputtype" in user_cfg:
        s = str(user_cfg["outputtype"]).lower()
        if s not in OUTPUTTYPES:
            raise ValueError("the specified output type is not supported")
        t = OUTPUTTYPES.index(s)
        cfg.outputtype = t - 2 if t >= 5 else t
    if "debuglevel" in user_cfg:
        s = str(user_cfg["debuglevel"]).upper()
        lvl = sum(1 << i for i, c in enumerate(DEBUGFLAGS) if c in s)
        if lvl == 0:
            raise ValueError("the specified debuglevel is not supported")
        cfg.debuglevel = lvl
    if "savedetflag" in user_cfg:
        v = user_cfg["savedetflag"]
        if isinstance(v, str):
            cfg.savedetflag = sum(1 << i for i, c in enumerate(SAVEFLAGS) if c in v.upper())
        else:
            cfg.savedetflag = int(v)
    if "srcpattern" in user_cfg:
        pat = np.asarray(user_cfg["srcpattern"], dtype=np.float32)
        cfg.srcpattern = ctypes.cast(_copy_to_c(pat.ravel(order="F"), ctypes.c_float), ctypes.c_void_p).value
    if "invcdf" in user_cfg or "angleinvcdf" in user_cfg:
        raise ValueError("user-defined inverse CDFs are not supported by this engine")
    if "shapes" in user_cfg:
        s = str(user_cfg["shapes"]).encode() + b"\0"
        p = _cmalloc(len(s) + 1)
        ctypes.memmove(p, s, len(s))
        cfg.shapedata = p
    if "bc" in user_cfg:
        s = str(user_cfg["bc"])
        if not s or len(s) > 12:
            raise ValueError("the 'bc' field must be a non-empty string with at most 12 characters")
        cfg.bc = s.encode()
    replay = None
    if "seed" in user_cfg:
        sv = user_cfg["seed"]
        if isinstance(sv, (int, np.integer, float, np.floating)):
            cfg.seed = int(sv)
        else:
            # src/pmcx.cpp:934-952: a [16, n] uint8 array of saved RNG states
            seeds = np.asfortranarray(np.asarray(sv, dtype=np.uint8))
            if seeds.ndim != 2 or seeds.shape[0] != SEEDBYTE:
                raise ValueError("the row number of cfg.seed does not match RNG seed byte-length")
            buf = seeds.ravel(order="F").copy()
            cfg.replay.seed = ctypes.cast(_copy_to_c(buf, ctypes.c_ubyte), ctypes.c_void_p).value
            cfg.seed = MCX_SEED_FROM_FILE
            cfg.nphoton = seeds.shape[1]
    if "detphotons" in user_cfg:
        d = user_cfg["detphotons"]
        if isinstance(d, dict):
            d = d.get("data")
        d = np.asfortranarray(np.asarray(d, dtype=np.float32))
        if d.ndim != 2:
            raise ValueError("the 'detphotons' field must be a 2D array [reclen, n]")
        replay = d
    if "gpuid" in user_cfg:
        g = user_cfg["gpuid"]
        cfg.deviceid = b"\0" * 256
        if isinstance(g, (int, np.integer)):
            if not (0 < int(g) < 256):
                raise ValueError("GPU id must be positive and can not be more than 256")
            cfg.gpuid = int(g)
            cfg.deviceid = b"0" * (int(g) - 1) + b"1"
        else:
            cfg.deviceid = str(g).encode()[:256]
    if "workload" in user_cfg:
        w = np.asarray(user_cfg["workload"], dtype=np.float32).ravel()
        for i, x in enumerate(w[:256]):
            cfg.workload[i] = float(x)
    return replay


def _reclen(cfg):
    f = cfg.savedetflag
    partial = (cfg.medianum - 1) * (((f >> 1) & 1) + ((f >> 2) & 1) + ((f >> 3) & 1))
    return partial + (f & 1) + 3 * (((f >> 4) & 1) + ((f >> 5) & 1)) + ((f >> 6) & 1) + 4 * ((f >> 7) & 1)


def run(cfg=None, **kwargs):
    """Run one simulation; cfg is a dict of pmcx fields (or keyword args)."""
    user_cfg = dict(cfg or {})
    user_cfg.update(kwargs)
    L = _lib.lib()
    c = Config()
    gpu = ctypes.POINTER(GPUInfo)()
    output = {}
    try:
        detps = _parse_config(user_cfg, c)
        ndev = L.mcx_list_gpu(ctypes.byref(c), ctypes.byref(gpu))
        _check()
        if not ndev:
            raise RuntimeError("No GPU device found")
        if detps is not None:
            # src/pmcx.cpp:1096: the baseline's detected photons drive the replay
            dims = (ctypes.c_int * 2)(detps.shape[0], detps.shape[1])
            flat = np.ascontiguousarray(detps.ravel(order="F"), dtype=np.float32)
            L.mcx_validatecfg(ctypes.byref(c), flat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), dims,
                              SEEDBYTE if c.seed == MCX_SEED_FROM_FILE else 0)
        else:
            L.mcx_validatecfg(ctypes.byref(c), None, None, SEEDBYTE if c.seed == MCX_SEED_FROM_FILE else 0)
        _check()
        if not c.vol or c.medianum == 0:
            raise ValueError("You must define 'vol' and 'prop' field.")
        ngates = int((c.tend - c.tstart) / c.tstep + 0.5)
        nvox = c.dim.x * c.dim.y * c.dim.z
        reclen = _reclen(c)
        # per-detector replay tallies carry one block of gates per detector;
        # RF replay adds an imaginary plane (src/pmcx.cpp:1115-1121)
        replay = c.seed == MCX_SEED_FROM_FILE
        blocks = c.detnum if (replay and c.replaydet == -1 and c.detnum > 1) else 1
        planes = 2 if (replay and c.outputtype in (OT_RF, OT_RFMUS)) else 1
        if c.issave2pt:
            c.exportfield = ctypes.cast(_libc.calloc(nvox * ngates * c.srcnum * blocks * planes, 4),
                                        ctypes.POINTER(ctypes.c_float))
        if c.issavedet:
            c.exportdetected = ctypes.cast(_cmalloc(reclen * c.maxdetphoton * 4), ctypes.POINTER(ctypes.c_float))
        if user_cfg.get("dumpmask", False):
            output["vol"] = np.ctypeslib.as_array(c.vol, shape=(nvox,)).copy().reshape(
                (c.dim.x, c.dim.y, c.dim.z), order="F")
        L.mcx_run_simulation(ctypes.byref(c), gpu)
        _check()
        if c.issavedet:
            n = int(c.detectedcount)
            if n > 0:
                d = np.ctypeslib.as_array(c.exportdetected, shape=(n * reclen,)).copy()
                output["detp"] = np.asfortranarray(d.reshape((n, reclen)).T)
        if c.issavedet and c.issaveseed:
            n = int(c.detectedcount)
            if n > 0 and c.seeddata:
                raw = np.ctypeslib.as_array(ctypes.cast(c.seeddata, ctypes.POINTER(ctypes.c_ubyte)), shape=(n * SEEDBYTE,))
                output["seeds"] = np.asfortranarray(raw.copy().reshape((n, SEEDBYTE)).T)
            else:
                output["seeds"] = np.zeros((SEEDBYTE, 0), dtype=np.uint8, order="F")
        if c.issave2pt:
            f = np.ctypeslib.as_array(c.exportfield, shape=(nvox * ngates * c.srcnum * blocks * planes,)).copy()
            # src/pmcx.cpp:1243-1262: [x, y, z, gates(, detectors)(, re/im)]
            d4 = c.detnum if (replay and c.replaydet == -1) else 1
            if planes > 1:
                shape = (c.srcnum * c.dim.x, c.dim.y, c.dim.z, ngates, d4, planes)
            else:
                shape = (c.srcnum * c.dim.x, c.dim.y, c.dim.z, ngates) + ((d4,) if d4 > 1 else ())
            if c.issaveref:
                # src/pmcx.cpp:1268-1289: void voxels hold the (negated)
                # escaping weight; split them out as 'dref'
                vol = np.ctypeslib.as_array(c.vol, shape=(nvox,)) & 0x7FFFFFFF
                g = f.reshape(-1, nvox)
                void = vol == 0
                dref = np.where(void[None, :], -g, 0.0).astype(np.float32)
                g[:, void] = 0.0
                output["dref"] = dref.reshape(-1).reshape(shape, order="F")
            output["flux"] = f.reshape(shape, order="F")
            output["stat"] = {
                "runtime": int(c.runtime),
                "nphoton": int(c.nphoton) * (c.respin if c.respin > 1 else 1),
                "energytot": float(c.energytot),
                "energyabs": float(c.energyabs),
                "normalizer": float(c.normalizer),
                "unitinmm": float(c.unitinmm),
                "workload": [float(c.workload[i]) for i in range(max(ndev, 1))],
            }
    finally:
        if gpu:
            L.mcx_cleargpuinfo(ctypes.byref(gpu))
        L.mcx_clearcfg(ctypes.byref(c))
    return output


def run_replicas(cfg, k, out=None):
    """k independent runs of one forward simulation (k a power of two, at
    most 64) in one launch of the single-simulation kernel; virtual thread
    idx belongs to run idx % k, so each run has nthread / k streams and
    about cfg['nphoton'] photons.  Returns {"cw": float32 [k, Nx, Ny, Nz]
    time-integrated flux (sum over gates x tstep, each run normalised by
    its own launched energy), "energytot", "energyabs": [k] FP64,
    "clear_ms", "kernel_ms", "reduce_ms" (device time of the tally clear,
    the transport kernel and the reduction), "wall_ms"}.  out: an optional
    device tensor of k * nvox float32 (x fastest) that receives the maps
    instead (then "cw" is absent)."""
    user_cfg = dict(cfg)
    user_cfg.setdefault("issavedet", 0)
    L = _lib.lib()
    c = Config()
    gpu = ctypes.POINTER(GPUInfo)()
    try:
        _parse_config(user_cfg, c)
        ndev = L.mcx_list_gpu(ctypes.byref(c), ctypes.byref(gpu))
        _check()
        if not ndev:
            raise RuntimeError("No GPU device found")
        L.mcx_validatecfg(ctypes.byref(c), None, None, 0)
        _check()
        nx, ny, nz = c.dim.x, c.dim.y, c.dim.z
        if out is not None:
            if out.numel() != int(k) * nx * ny * nz or not out.is_cuda:
                raise ValueError("out must be a device tensor of k * nvox floats")
            cw = None
            dst = ctypes.cast(ctypes.c_void_p(out.data_ptr()), ctypes.POINTER(ctypes.c_float))
        else:
            cw = np.empty((int(k), nz, ny, nx), dtype=np.float32)
            dst = cw.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        energy = np.zeros(2 * int(k), dtype=np.float64)
        stats = np.zeros(4, dtype=np.float64)
        rc = L.mcx_amd_run_replicas(ctypes.byref(c), gpu, int(k), dst,
                                    energy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    stats.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc != 0:
            raise RuntimeError("mcx_amd_run_replicas failed: %s" % (_lib.last_error(),))
    finally:
        if gpu:
            L.mcx_cleargpuinfo(ctypes.byref(gpu))
        L.mcx_clearcfg(ctypes.byref(c))
    res = {"energytot": energy[1::2].copy(), "energyabs": energy[1::2] - energy[0::2], "clear_ms": float(stats[0]),
           "kernel_ms": float(stats[1]), "reduce_ms": float(stats[2]), "wall_ms": float(stats[3])}
    if cw is not None:
        res["cw"] = np.ascontiguousarray(cw.transpose(0, 3, 2, 1))
    return res


def gpuinfo():
    """List HIP devices as dicts (src/pmcx.cpp:1404-1439)."""
    L = _lib.lib()
    c = Config()
    L.mcx_initcfg(ctypes.byref(c))
    c.parentid = mpPython
    c.flog = None
    c.isgpuinfo = 3
    gpu = ctypes.POINTER(GPUInfo)()
    out = []
    n = L.mcx_list_gpu(ctypes.byref(c), ctypes.byref(gpu))
    try:
        if n and gpu:
            for i in range(gpu[0].devcount):
                g = gpu[i]
                out.append({"name": g.name.decode(errors="replace"), "id": g.id, "devcount": g.devcount,
                            "major": g.major, "minor": g.minor, "globalmem": g.globalmem, "constmem": g.constmem,
                            "sharedmem": g.sharedmem, "regcount": g.regcount, "clock": g.clock, "sm": g.sm,
                            "core": g.core, "autoblock": g.autoblock, "autothread": g.autothread,
                            "maxgate": g.maxgate})
    finally:
        if gpu:
            L.mcx_cleargpuinfo(ctypes.byref(gpu))
    return out


def version():
    return "v2025.10-mi355x"
