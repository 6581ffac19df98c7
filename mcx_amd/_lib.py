"""ctypes binding of libmcx_amd.so (the C-ABI boundary).

The Structure classes mirror include/mcx_abi.h field for field; the library
exports mcx_amd_abi_layout() so the mirror is checked against the compiled
layout at load time.  Loading fails loudly when the shared library is
missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcx_amd.so")

MAX_SESSION_LENGTH = 256
MAX_PATH_LENGTH = 1024
MAX_DEVICE = 256
MAX_LANG_ID = 32
mpStandalone, mpMATLAB, mpPython = 0, 1, 2


class Float3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class Float4(ctypes.Structure):
    # 16-byte alignment of the C type is reproduced with explicit padding
    # fields in Config (ctypes cannot raise a structure's alignment).
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float), ("w", ctypes.c_float)]


class Uint3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint), ("y", ctypes.c_uint), ("z", ctypes.c_uint)]


class Medium(ctypes.Structure):
    _fields_ = [("mua", ctypes.c_float), ("mus", ctypes.c_float), ("g", ctypes.c_float), ("n", ctypes.c_float)]


class History(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_char * 4), ("version", ctypes.c_uint), ("maxmedia", ctypes.c_uint),
        ("detnum", ctypes.c_uint), ("colcount", ctypes.c_uint), ("totalphoton", ctypes.c_uint),
        ("detected", ctypes.c_uint), ("savedphoton", ctypes.c_uint), ("unitinmm", ctypes.c_float),
        ("seedbyte", ctypes.c_uint), ("normalizer", ctypes.c_float), ("respin", ctypes.c_int),
        ("srcnum", ctypes.c_uint), ("savedetflag", ctypes.c_uint), ("totalsource", ctypes.c_uint),
        ("reserved", ctypes.c_int * 1),
    ]


class Replay(ctypes.Structure):
    _fields_ = [("detid", ctypes.c_void_p), ("seed", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("tof", ctypes.c_void_p)]


class GPUInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * MAX_SESSION_LENGTH), ("id", ctypes.c_int), ("devcount", ctypes.c_int),
        ("major", ctypes.c_int), ("minor", ctypes.c_int), ("globalmem", ctypes.c_size_t),
        ("constmem", ctypes.c_size_t), ("sharedmem", ctypes.c_size_t), ("regcount", ctypes.c_int),
        ("clock", ctypes.c_int), ("sm", ctypes.c_int), ("core", ctypes.c_int), ("autoblock", ctypes.c_int),
        ("autothread", ctypes.c_int), ("maxgate", ctypes.c_uint), ("maxmpthread", ctypes.c_int),
    ]


class Config(ctypes.Structure):
    """Mirror of Config (src/mcx_utils.h:172-285)."""
    _fields_ = [
        ("nphoton", ctypes.c_size_t), ("nblocksize", ctypes.c_uint), ("nthread", ctypes.c_uint),
        ("seed", ctypes.c_int), ("_pad0", ctypes.c_char * 12),
        ("srcpos", Float4), ("srcdir", Float4), ("srciquv", Float4),
        ("tstart", ctypes.c_float), ("tstep", ctypes.c_float), ("tend", ctypes.c_float),
        ("steps", Float3), ("dim", Uint3), ("crop0", Uint3), ("crop1", Uint3),
        ("medianum", ctypes.c_uint), ("polmedianum", ctypes.c_uint), ("detnum", ctypes.c_uint),
        ("maxdetphoton", ctypes.c_uint), ("detradius", ctypes.c_float), ("sradius", ctypes.c_float),
        ("prop", ctypes.POINTER(Medium)), ("polprop", ctypes.c_void_p), ("detpos", ctypes.POINTER(Float4)),
        ("smatrix", ctypes.c_void_p), ("maxgate", ctypes.c_uint), ("respin", ctypes.c_int),
        ("printnum", ctypes.c_int), ("gpuid", ctypes.c_int), ("vol", ctypes.POINTER(ctypes.c_uint)),
        ("session", ctypes.c_char * MAX_SESSION_LENGTH),
        ("isrowmajor", ctypes.c_byte), ("isreflect", ctypes.c_byte), ("isref3", ctypes.c_byte),
        ("isrefint", ctypes.c_byte), ("isnormalized", ctypes.c_byte), ("issavedet", ctypes.c_byte),
        ("issave2pt", ctypes.c_byte), ("isgpuinfo", ctypes.c_byte), ("isspecular", ctypes.c_byte),
        ("issrcfrom0", ctypes.c_byte), ("isdumpmask", ctypes.c_byte), ("autopilot", ctypes.c_byte),
        ("issaveseed", ctypes.c_byte), ("issaveexit", ctypes.c_byte), ("issaveref", ctypes.c_byte),
        ("ismomentum", ctypes.c_byte), ("istrajstokes", ctypes.c_byte), ("isdumpjson", ctypes.c_byte),
        ("internalsrc", ctypes.c_byte),
        ("zipid", ctypes.c_int), ("srctype", ctypes.c_byte), ("outputtype", ctypes.c_byte),
        ("outputformat", ctypes.c_byte), ("faststep", ctypes.c_byte),
        ("minenergy", ctypes.c_float), ("unitinmm", ctypes.c_float), ("omega", ctypes.c_float),
        ("lambda_", ctypes.c_float), ("flog", ctypes.c_void_p), ("his", History),
        ("exportfield", ctypes.POINTER(ctypes.c_float)), ("exportdetected", ctypes.POINTER(ctypes.c_float)),
        ("detectedcount", ctypes.c_ulong), ("rootpath", ctypes.c_char * MAX_PATH_LENGTH),
        ("shapedata", ctypes.c_void_p), ("extrajson", ctypes.c_void_p),
        ("maxvoidstep", ctypes.c_int), ("voidtime", ctypes.c_int), ("_pad1", ctypes.c_char * 8),
        ("srcparam1", Float4), ("srcparam2", Float4), ("srcnum", ctypes.c_uint),
        ("srcpattern", ctypes.c_void_p), ("replay", Replay), ("seeddata", ctypes.c_void_p),
        ("replaydet", ctypes.c_int), ("seedfile", ctypes.c_char * MAX_PATH_LENGTH),
        ("jsonfile", ctypes.c_char * MAX_PATH_LENGTH), ("debuglevel", ctypes.c_uint),
        ("savedetflag", ctypes.c_uint), ("deviceid", ctypes.c_char * MAX_DEVICE),
        ("workload", ctypes.c_float * MAX_DEVICE), ("parentid", ctypes.c_int), ("runtime", ctypes.c_uint),
        ("langid", ctypes.c_char * MAX_LANG_ID),
        ("energytot", ctypes.c_double), ("energyabs", ctypes.c_double), ("energyesc", ctypes.c_double),
        ("normalizer", ctypes.c_float), ("maxjumpdebug", ctypes.c_uint), ("debugdatalen", ctypes.c_uint),
        ("gscatter", ctypes.c_uint), ("exportdebugdata", ctypes.c_void_p), ("mediabyte", ctypes.c_uint),
        ("dx", ctypes.c_void_p), ("dy", ctypes.c_void_p), ("dz", ctypes.c_void_p),
        ("bc", ctypes.c_char * 13), ("nphase", ctypes.c_uint), ("invcdf", ctypes.c_void_p),
        ("nangle", ctypes.c_uint), ("angleinvcdf", ctypes.c_void_p), ("srcid", ctypes.c_int),
        ("extrasrclen", ctypes.c_uint), ("srcdata", ctypes.c_void_p),
    ]


# every symbol include/mcx_abi.h and include/mcx_amd_ext.h declare
ABI_SYMBOLS = [
    "mcx_run_simulation", "mcx_list_gpu", "mcx_initcfg", "mcx_clearcfg", "mcx_preprocess", "mcx_maskdet",
    "mcx_validatecfg", "mcx_cleargpuinfo", "mcx_error", "mcx_flush", "mcx_normalize", "mcx_readconfig",
    "mcx_run_from_json", "mcx_savedata", "mcx_savedetphoton", "mcx_float2half2", "mcx_last_error", "mcx_clear_error",
    "mcx_rng_test", "mcx_amd_set_tally_mode", "mcx_amd_set_counters", "mcx_amd_set_hot_box", "mcx_amd_set_replay_index",
    "mcx_amd_set_tally_blocks", "mcx_amd_set_refill", "mcx_amd_seed_words",
    "mcx_amd_run_replicas",
    "mcx_amd_last_stats",
    "mcx_amd_abi_layout", "mcx_amd_atomic_bench", "mcx_run_batch",
    "mcx_dn_conv3_fwd", "mcx_dn_conv3_fwd_cat", "mcx_dn_weight_flip", "mcx_dn_conv3_wgrad", "mcx_dn_wgrad_reduce", "mcx_dn_bias_grad",
    "mcx_dn_bias_grad_noalloc",
    "mcx_dn_maxpool_fwd", "mcx_dn_maxpool_bwd", "mcx_dn_upsample_fwd", "mcx_dn_upsample_bwd", "mcx_dn_add", "mcx_dn_relu_mask", "mcx_dn_roi_max", "mcx_dn_roi_counts",
    "mcx_dn_head_fwd", "mcx_dn_head_bwd", "mcx_dn_adam", "mcx_dn_f32_to_bf16", "mcx_dn_features",
    "mcx_dn_inv_features",
]

_lib = None


def _setup(L):
    P = ctypes.POINTER
    L.mcx_initcfg.argtypes = [P(Config)]
    L.mcx_clearcfg.argtypes = [P(Config)]
    L.mcx_preprocess.argtypes = [P(Config)]
    L.mcx_maskdet.argtypes = [P(Config)]
    L.mcx_validatecfg.argtypes = [P(Config), P(ctypes.c_float), P(ctypes.c_int), ctypes.c_int]
    L.mcx_list_gpu.argtypes = [P(Config), P(P(GPUInfo))]
    L.mcx_list_gpu.restype = ctypes.c_int
    L.mcx_run_simulation.argtypes = [P(Config), P(GPUInfo)]
    L.mcx_cleargpuinfo.argtypes = [P(P(GPUInfo))]
    L.mcx_readconfig.argtypes = [ctypes.c_char_p, P(Config)]
    L.mcx_run_from_json.argtypes = [ctypes.c_char_p]
    L.mcx_run_from_json.restype = ctypes.c_int
    L.mcx_savedata.argtypes = [P(ctypes.c_float), ctypes.c_size_t, P(Config)]
    L.mcx_savedetphoton.argtypes = [P(ctypes.c_float), ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P(Config)]
    L.mcx_normalize.argtypes = [P(ctypes.c_float), ctypes.c_float, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int]
    L.mcx_float2half2.argtypes = [P(ctypes.c_float)]
    L.mcx_float2half2.restype = ctypes.c_int
    L.mcx_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.mcx_last_error.restype = ctypes.c_int
    L.mcx_rng_test.argtypes = [ctypes.c_uint, P(ctypes.c_float), ctypes.c_uint]
    L.mcx_rng_test.restype = ctypes.c_int
    L.mcx_amd_set_tally_mode.argtypes = [ctypes.c_int]
    L.mcx_amd_set_tally_mode.restype = ctypes.c_int
    L.mcx_amd_set_counters.argtypes = [ctypes.c_int]
    L.mcx_amd_set_hot_box.argtypes = [ctypes.c_int]
    L.mcx_amd_set_hot_box.restype = ctypes.c_int
    L.mcx_amd_run_replicas.argtypes = [P(Config), P(GPUInfo), ctypes.c_int, P(ctypes.c_float), P(ctypes.c_double),
                                       P(ctypes.c_double)]
    L.mcx_amd_run_replicas.restype = ctypes.c_int
    L.mcx_amd_set_refill.argtypes = [ctypes.c_int]
    L.mcx_amd_set_refill.restype = ctypes.c_int
    L.mcx_amd_seed_words.argtypes = [ctypes.c_uint, P(ctypes.c_uint), P(ctypes.c_ulonglong), ctypes.c_int]
    L.mcx_amd_seed_words.restype = ctypes.c_int
    L.mcx_amd_set_tally_blocks.argtypes = [ctypes.c_int]
    L.mcx_amd_set_tally_blocks.restype = ctypes.c_int
    L.mcx_amd_set_replay_index.argtypes = [ctypes.c_int]
    L.mcx_amd_set_replay_index.restype = ctypes.c_int
    L.mcx_amd_last_stats.argtypes = [P(ctypes.c_double), ctypes.c_int]
    L.mcx_amd_last_stats.restype = ctypes.c_int
    L.mcx_amd_abi_layout.argtypes = [P(ctypes.c_size_t), ctypes.c_int]
    L.mcx_amd_atomic_bench.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint, P(ctypes.c_double)]
    L.mcx_amd_atomic_bench.restype = ctypes.c_int
    L.mcx_amd_abi_layout.restype = ctypes.c_int
    L.mcx_run_batch.argtypes = [P(P(Config)), ctypes.c_int, P(GPUInfo), P(McxBatchOptions)]
    L.mcx_run_batch.restype = ctypes.c_int
    V, I, LL, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
    dn = {
        "mcx_dn_conv3_fwd": [V, I, V, I, V, I, V, V, I, I, I, I, I, I, I, I, V],
        "mcx_dn_conv3_fwd_cat": [V, I, I, V, I, V, I, V, V, I, I, I, I, I, I, I, I, V],
        "mcx_dn_weight_flip": [V, I, V, I, I, I, V],
        "mcx_dn_conv3_wgrad": [V, I, V, I, V, I, V, I, I, I, I, I, I, I, V],
        "mcx_dn_wgrad_reduce": [V, I, V, I, I, I, V],
        "mcx_dn_bias_grad": [V, I, V, I, V, I, V, LL, I, V],
        "mcx_dn_bias_grad_noalloc": [V, I, V, I, I, V, LL, I, V],
        "mcx_dn_maxpool_fwd": [V, I, V, I, V, I, I, I, I, I, V],
        "mcx_dn_maxpool_bwd": [V, I, V, V, I, I, I, I, I, I, V],
        "mcx_dn_upsample_fwd": [V, I, V, I, I, I, I, I, I, V],
        "mcx_dn_upsample_bwd": [V, I, V, I, I, I, I, I, I, V],
        "mcx_dn_add": [V, I, V, I, V, I, LL, I, V],
        "mcx_dn_relu_mask": [V, I, V, I, LL, I, V],
        "mcx_dn_roi_max": [V, I, LL, V, V],
        "mcx_dn_roi_counts": [V, I, I, I, I, V, V, V, V],
        "mcx_dn_head_fwd": [V, I, V, V, V, I, V, V, V, V, I, LL, I, V],
        "mcx_dn_head_bwd": [V, I, V, V, V, I, V, I, V, LL, I, V],
        "mcx_dn_adam": [V, V, V, V, V, LL, F, F, F, F, I, V],
        "mcx_dn_f32_to_bf16": [V, V, LL, V],
        "mcx_dn_features": [V, V, V, I, V, I, I, I, I, F, F, F, V],
        "mcx_dn_inv_features": [V, V, I, I, I, I, F, F, F, V],
    }
    for name, args in dn.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int


class McxBatchOptions(ctypes.Structure):
    """Mirror of McxBatchOptions in include/mcx_amd_ext.h."""
    _fields_ = [
        ("tally_mode", ctypes.c_int), ("photons_per_lane", ctypes.c_int), ("block_size", ctypes.c_int),
        ("use_graph", ctypes.c_int), ("repeat", ctypes.c_int), ("download", ctypes.c_int),
        ("tune_fraction", ctypes.c_float), ("chosen_photons_per_lane", ctypes.c_int),
        ("chosen_block_size", ctypes.c_int), ("vgprs", ctypes.c_int), ("kernel_ms", ctypes.c_float),
        ("total_ms", ctypes.c_float), ("photons", ctypes.c_double), ("lanes", ctypes.c_double),
        ("counters", ctypes.c_double * 4), ("tally_cache_cells", ctypes.c_int),
        ("chosen_tally_cache_cells", ctypes.c_int), ("wave_iterations", ctypes.c_double),
    ]


def check_layout(L):
    """Compare the ctypes mirror with the compiled struct layout."""
    out = (ctypes.c_size_t * 16)()
    n = L.mcx_amd_abi_layout(out, 16)
    want = [ctypes.sizeof(Config), ctypes.sizeof(GPUInfo), ctypes.sizeof(History), Config.srcpos.offset,
            Config.session.offset, Config.flog.offset, Config.his.offset, Config.exportfield.offset,
            Config.srcparam1.offset, Config.deviceid.offset, Config.energytot.offset, Config.bc.offset,
            Config.srcdata.offset, ctypes.sizeof(McxBatchOptions), McxBatchOptions.photons.offset]
    got = list(out[:n])
    if got != want[:n]:
        raise RuntimeError("Config layout mismatch between libmcx_amd.so %s and the ctypes mirror %s" % (got, want))


def _share_torch_hip_runtime():
    """PyTorch-ROCm ships its own HIP runtime under ROCm's soname
    (libamdhip64.so.7) but links it by the name libamdhip64.so.  If this
    library loaded ROCm's runtime first, a later `import torch` would load a
    second HIP and HSA runtime into the process, which then finds no
    device.  Loading torch's runtime first makes this library bind to it,
    so both share one runtime (and can share device pointers and streams)."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def lib():
    """Load libmcx_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmcx_amd.so is not built; run __graft_entry__.build() or make -C mcx_amd/csrc")
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        _setup(L)
        check_layout(L)
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    code = lib().mcx_last_error(buf, 1024)
    return code, buf.value.decode(errors="replace")


def atomic_bench(mode, nelem, nper, nblocks):
    """Scattered-atomic rate microbenchmark (see include/mcx_amd_ext.h)."""
    out = (ctypes.c_double * 11)()
    rc = lib().mcx_amd_atomic_bench(mode, nelem, nper, nblocks, out)
    if rc != 0:
        raise RuntimeError("atomic bench failed (%d)" % rc)
    return {"atomics_per_s": out[0], "ms": out[1], "sum_ratio": out[2], "xcc_workgroups": list(out[3:11])}


def seed_words(seed, counts):
    """The host seeding's per-thread seed words: consecutive draws of
    counts[i] words from srand(seed), back to back (uint32 array)."""
    import numpy as np
    counts = [int(c) for c in counts]
    out = np.zeros(max(1, sum(counts)), dtype=np.uint32)
    cc = (ctypes.c_ulonglong * max(1, len(counts)))(*counts)
    rc = lib().mcx_amd_seed_words(int(seed), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), cc, len(counts))
    if rc != 0:
        raise RuntimeError("mcx_amd_seed_words failed (%d)" % rc)
    return out[:sum(counts)]


def last_stats():
    out = (ctypes.c_double * 18)()
    n = lib().mcx_amd_last_stats(out, 18)
    keys = ["kernel_ms", "wall_ms", "nstep", "ndeposit", "nscatter", "nlaunch", "vgprs", "nthread", "nblock",
            "nlaunches", "errlanes", "hot_cells", "lds_bytes", "ndeposit_hot", "tally_blocks",
            "tally_requests", "wave_iterations", "lanes"]
    return dict(zip(keys, list(out[:n])))
