"""mcx_amd -- MI355X-native Monte Carlo photon transport (drop-in for MCX's
mcx_run_simulation path).

The engine is libmcx_amd.so (HIP kernel + C++ host layer, C ABI declared in
include/mcx_abi.h).  This package exposes the pmcx-style dictionary API.
"""
from .pmcx import gpuinfo, run, run_replicas, version  # noqa: F401
from .batch import run_batch  # noqa: F401
from . import _lib  # noqa: F401


def set_tally_mode(mode):
    """0: FP32 + MAX_ACCUM shadow (reference), 1: FP32 hi/lo compensated, 2: FP64,
    3: FP64 with one private copy per XCD."""
    names = {"f32": 0, "shadow": 0, "hilo": 1, "f64": 2}
    m = names.get(mode, mode) if isinstance(mode, str) else int(mode)
    old = _lib.lib().mcx_amd_set_tally_mode(m)
    if old < 0:
        raise ValueError("invalid tally mode %r" % (mode,))
    return old


def set_counters(enable):
    _lib.lib().mcx_amd_set_counters(1 if enable else 0)


def set_hot_box(enable):
    """Sum the tally cells next to the source in each workgroup's LDS
    (default on; FP64 and hi/lo tallies).  Returns the previous setting."""
    return bool(_lib.lib().mcx_amd_set_hot_box(1 if enable else 0))


def set_tally_blocks(mode):
    """FP64 tally in 2x2x2 voxel blocks, summed per lane in LDS and added
    to the tally 8 lanes to a block: False/0 off, True/1 on, 2 or "auto"
    (default) on when the medium scatters at most twice per voxel length.
    Returns the previous mode."""
    m = 2 if mode == "auto" else int(mode)
    old = _lib.lib().mcx_amd_set_tally_blocks(m)
    if old < 0:
        raise ValueError("invalid tally block mode %r" % (mode,))
    return old


def set_refill(mode):
    """Autopilot lanes run several virtual threads each, taking the next
    unstarted one when theirs is done.  0/False: off; 1/True (default): at
    least 64 photons per virtual thread, at most 8 virtual threads per lane;
    2 or "fast": about one photon per virtual thread, at most 32 per lane
    (shortest tail; each stream's biased first random number then weighs on
    every photon or two).  Returns the previous mode."""
    m = 2 if mode == "fast" else int(mode)
    old = _lib.lib().mcx_amd_set_refill(m)
    if old < 0:
        raise ValueError("invalid refill mode %r" % (mode,))
    return old


def set_replay_index(reference):
    """Number replayed photons as the reference does (True) or contiguously
    (False, default).  Returns the previous setting."""
    return bool(_lib.lib().mcx_amd_set_replay_index(1 if reference else 0))


def last_stats():
    return _lib.last_stats()
