"""Helpers shared by the parity tests: run the same pmcx dict through the HIP
engine (C ABI via mcx_amd.run) and through the CPU oracle."""
import numpy as np

import oracle_py

SAVEFLAGS = "DSPMXVWI"


def flag_bits(v):
    if isinstance(v, str):
        return sum(1 << i for i, c in enumerate(SAVEFLAGS) if c in v.upper())
    return int(v)


def oracle_run(cfg, nthread, **extra):
    """Run the oracle on a pmcx dict whose positions are 0-based (issrcfrom0=1)."""
    assert cfg.get("issrcfrom0", 0) == 1
    srctypes = ["pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier", "arcsine", "disk", "fourierx",
                "fourierx2d", "zgaussian", "line", "slit", "pencilarray", "pattern3d", "hyperboloid", "ring"]
    st = cfg.get("srctype", "pencil")
    srctype = srctypes.index(st) if isinstance(st, str) else int(st)
    det = cfg.get("detpos")
    issavedet = int(cfg.get("issavedet", 1 if det is not None else 0))
    ot = {"flux": 0, "fluence": 1, "energy": 2, "jacobian": 3, "wl": 3, "wp": 4, "nscat": 4, "wm": 5, "rf": 6,
          "length": 7, "rfmus": 8, "wltof": 9, "wptof": 10}[cfg.get("outputtype", "flux")]
    kw = dict(
        tstart=cfg["tstart"], tend=cfg["tend"], tstep=cfg["tstep"], srctype=srctype,
        srcparam1=cfg.get("srcparam1", (0, 0, 0, 0)), srcparam2=cfg.get("srcparam2", (0, 0, 0, 0)),
        detpos=None if det is None else np.asarray(det, dtype=np.float32),
        isreflect=int(cfg.get("isreflect", 1)), isspecular=int(cfg.get("isspecular", 0)),
        unitinmm=float(cfg.get("unitinmm", 1.0)), outputtype=ot,
        savedetflag=flag_bits(cfg.get("savedetflag", 0x5)), issavedet=issavedet,
        maxdetphoton=int(cfg.get("maxdetphoton", 1000000)), minenergy=float(cfg.get("minenergy", 0.0)),
        bc=cfg.get("bc", ""),
        srcpattern=cfg.get("srcpattern"),
        issaveref=int(cfg.get("issaveref", 0)),
        omega=float(cfg.get("omega", 0.0)),
    )
    vol = cfg["vol"]
    if np.asarray(vol).ndim == 4 or int(cfg.get("mediabyte", 0)) > 4:
        # property-word media: the words the engine receives (mcx_amd.media,
        # pinned by tests/test_media.py)
        from mcx_amd import media
        if int(cfg.get("mediabyte", 0)) > 4:
            words, mb = np.asarray(vol, dtype=np.uint32), int(cfg["mediabyte"])
            dims = words.shape
        else:
            words, mb, dims = media.encode_volume(vol, float(cfg.get("unitinmm", 1.0)))
        nvox = int(np.prod(dims))
        flat = np.asarray(words, dtype=np.uint32).ravel(order="F")
        if mb == media.MEDIA_2LABEL_SPLIT:
            flat = media.svmc_planes(flat)
        vol = flat[:nvox].reshape(dims, order="F")
        kw["mediaformat"] = mb
        if mb in (media.MEDIA_ASGN_F2H, media.MEDIA_2LABEL_SPLIT):
            kw["words2"] = flat[nvox:].reshape(dims, order="F")
    kw.update(extra)
    return oracle_py.run(vol, cfg["prop"], cfg["srcpos"], list(cfg["srcdir"]), cfg["nphoton"], nthread,
                         cfg["seed"], **kw)


def seq_sum(x):
    """Left-to-right double sum, the order the host layer uses."""
    s = 0.0
    for v in np.asarray(x, dtype=np.float64):
        s += v
    return s


def normalize_like_engine(field, energytot, tstep, unitinmm=1.0, outputtype=0):
    """Normalisation of src/mcx_core.cu:3823-3925 applied to an FP64 tally."""
    vvox = unitinmm ** 3
    if outputtype in (0, 1):
        scale = np.float32(unitinmm / (energytot * vvox * tstep))
        if outputtype == 1:
            scale = np.float32(scale * np.float32(tstep))
    else:
        scale = np.float32(1.0 / energytot)
    return field * float(scale), float(scale)


def sort_rows(a):
    a = np.asarray(a)
    if a.size == 0:
        return a
    idx = np.lexsort(a.T[::-1])
    return a[idx]
