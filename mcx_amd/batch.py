"""Batched execution: many independent simulations in one persistent launch.

run_batch(cfgs) takes a list of pmcx dicts (as accepted by run()) and calls
mcx_run_batch through the C ABI (include/mcx_amd_ext.h).  Each result is the
same dict run() returns for that configuration ({"flux", "stat"}); the second
return value describes the launch (photons, lanes, chosen shape, times).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import Config, GPUInfo, McxBatchOptions
from .pmcx import _check, _parse_config


def _check_host_memory(cfgs, mode):
    """Refuse a download whose arrays would not fit in 80 % of the host
    memory available now (the arrays are filled by the device copies)."""
    if not mode:
        return
    need = 0
    for c in cfgs:
        nvox = int(np.prod(np.asarray(c["vol"]).shape[:3]))
        gates = int((c["tend"] - c["tstart"]) / c["tstep"] + 0.5) if mode == 1 else 1
        need += 4 * nvox * max(gates, 1)
    try:
        avail = os.sysconf("SC_AVPHYS_PAGES") * os.sysconf("SC_PAGE_SIZE")
    except (ValueError, OSError):
        return
    if need > 0.8 * avail:
        raise RuntimeError("run_batch: the %s of %d instances needs %.1f GB of host memory, %.1f GB available"
                           % ("fluence" if mode == 1 else "time-integrated maps", len(cfgs), need / 1e9, avail / 1e9))


def run_batch(cfgs, tally=1, photons_per_lane=0, block_size=0, use_graph=False, repeat=1, download=True,
              tune_fraction=0.0, tally_cache=0):
    """download: True for each instance's fluence per gate ("flux",
    [x,y,z,gate]), "cw" for the time-integrated map reduced on the device
    ("cw", [x,y,z]: tstep x the sum over gates), False for none.  The
    arrays are numpy-owned; the library writes into them directly.
    tally_cache: per-wave LDS tally cache cells (< 0 none, > 0 at most
    that many, 0 the autotuner's choice or the largest that fits)."""
    mode = 2 if download == "cw" else (1 if download else 0)
    L = _lib.lib()
    _check_host_memory(cfgs, mode)
    n = len(cfgs)
    if n == 0:
        raise ValueError("run_batch needs at least one configuration")
    cs = [Config() for _ in range(n)]
    gpu = ctypes.POINTER(GPUInfo)()
    results = []
    outs = []
    try:
        for i, (user_cfg, c) in enumerate(zip(cfgs, cs)):
            _parse_config(dict(user_cfg), c)
            if i == 0:
                ndev = L.mcx_list_gpu(ctypes.byref(c), ctypes.byref(gpu))
                _check()
                if not ndev:
                    raise RuntimeError("No GPU device found")
            else:
                c.deviceid = cs[0].deviceid
            L.mcx_validatecfg(ctypes.byref(c), None, None, 0)
            _check()
            if not c.vol or c.medianum == 0:
                raise ValueError("You must define 'vol' and 'prop' field.")
            if mode:
                ngates = int((c.tend - c.tstart) / c.tstep + 0.5)
                nvox = c.dim.x * c.dim.y * c.dim.z
                buf = np.empty(nvox * (ngates if mode == 1 else 1), dtype=np.float32)
                outs.append(buf)
                c.exportfield = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        arr = (ctypes.POINTER(Config) * n)(*[ctypes.pointer(c) for c in cs])
        opt = McxBatchOptions()
        opt.tally_mode = int(tally)
        opt.photons_per_lane = int(photons_per_lane)
        opt.block_size = int(block_size)
        opt.use_graph = 1 if use_graph else 0
        opt.repeat = int(repeat)
        opt.download = mode
        opt.tune_fraction = float(tune_fraction)
        opt.tally_cache_cells = int(tally_cache)
        rc = L.mcx_run_batch(arr, n, gpu, ctypes.byref(opt))
        if rc != 0:
            _check()
            raise RuntimeError("mcx_run_batch failed (%d)" % rc)
        for i, c in enumerate(cs):
            ngates = int((c.tend - c.tstart) / c.tstep + 0.5)
            out = {"stat": {"runtime": int(c.runtime), "nphoton": int(c.nphoton), "energytot": float(c.energytot),
                            "energyabs": float(c.energyabs), "normalizer": float(c.normalizer),
                            "unitinmm": float(c.unitinmm)}}
            if mode == 1:
                out["flux"] = outs[i].reshape((c.dim.x, c.dim.y, c.dim.z, ngates), order="F")
            elif mode == 2:
                out["cw"] = outs[i].reshape((c.dim.x, c.dim.y, c.dim.z), order="F")
            results.append(out)
        info = {"photons": opt.photons, "lanes": int(opt.lanes), "kernel_ms": float(opt.kernel_ms),
                "total_ms": float(opt.total_ms), "photons_per_lane": int(opt.chosen_photons_per_lane),
                "block_size": int(opt.chosen_block_size), "vgprs": int(opt.vgprs),
                "tally_cache_cells": int(opt.chosen_tally_cache_cells),
                "counters": dict(zip(("nstep", "ndeposit", "nscatter", "nlaunch"), list(opt.counters)),
                                 wave_iterations=float(opt.wave_iterations))}
    finally:
        if gpu:
            L.mcx_cleargpuinfo(ctypes.byref(gpu))
        for c in cs:
            if mode:
                c.exportfield = None  # numpy-owned
            L.mcx_clearcfg(ctypes.byref(c))
    return results, info
