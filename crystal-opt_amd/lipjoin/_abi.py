"""ctypes mirror of include/lipjoin.h (the C-ABI of liblipjoin.so).

The structures below must match the C layout byte for byte; the library
reports its own struct sizes (lj_abi_struct_size) and `check_layout()`
compares them.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liblipjoin.so")

MAX_DIMS = 4
MAX_PRED_STEPS = 4

# lj_status
OK, EINVAL, ENOMEM, EBUDGET, EHIP, EUNSUPPORTED, EDOMAIN, ETIMEOUT, ESTATE = range(9)
STATUS_NAMES = {
    0: "LJ_OK", 1: "LJ_EINVAL", 2: "LJ_ENOMEM", 3: "LJ_EBUDGET", 4: "LJ_EHIP",
    5: "LJ_EUNSUPPORTED", 6: "LJ_EDOMAIN", 7: "LJ_ETIMEOUT", 8: "LJ_ESTATE",
}
# predicates
COMB_SET, COMB_AND, COMB_OR = 0, 1, 2
CMP_LT, CMP_GT, CMP_LTE, CMP_GTE, CMP_EQ = 0, 1, 2, 3, 4
# tables / filters / methods
HT_PHT2, HT_PHT1, HT_TEXTBOOK = 0, 1, 2
BLOOM_NONE, BLOOM_MOD, BLOOM_MULT = 0, 1, 2
BLOOM_STANDARD, BLOOM_BLOCKED = 0, 1
METHOD_NAIVE, METHOD_LIP, METHOD_LIP_ADAPTIVE = 0, 1, 2
# generators
GEN_UNIFORM, GEN_DATEKEY, GEN_SEQUENCE, GEN_YEAR_OF_DAY, GEN_DATEKEY_SEQ = 0, 1, 2, 3, 4
JOIN_SUM_ONLY = 1        # lj_join_desc.flags
TUNE_SEMI_JOIN = 1       # lj_tune_bloom flags


class PredStep(C.Structure):
    _fields_ = [("comb", C.c_int32), ("cmp", C.c_int32), ("value", C.c_int32)]


class Pred(C.Structure):
    _fields_ = [("n_steps", C.c_int32), ("steps", PredStep * MAX_PRED_STEPS)]


class DimDesc(C.Structure):
    _fields_ = [
        ("keys", C.c_void_p), ("attr", C.c_void_p), ("payload", C.c_void_p),
        ("n_rows", C.c_int64), ("pred", Pred), ("ht_kind", C.c_int32),
        ("keys_min", C.c_int32), ("ht_len", C.c_int64), ("bloom_hash", C.c_int32),
        ("bloom_layout", C.c_int32), ("bloom_k", C.c_int32), ("reserved0", C.c_int32),
        ("bloom_bits", C.c_int64),
    ]


class DimInfo(C.Structure):
    _fields_ = [
        ("n_selected", C.c_int64), ("n_oob", C.c_int64), ("ht_words", C.c_int64),
        ("bloom_words", C.c_int64), ("bloom_bits_set", C.c_int64), ("time_build_ms", C.c_float),
    ]


class ProbeStats(C.Structure):
    _fields_ = [
        ("rows_in", C.c_int64 * (MAX_DIMS + 1)), ("lines_in", C.c_int64 * (MAX_DIMS + 1)),
        ("rows_bloom_pass", C.c_int64), ("sample_rows", C.c_int64), ("sample_lines", C.c_int64),
        ("order_used", C.c_int32 * MAX_DIMS), ("n_oob", C.c_int64),
    ]


class JoinDesc(C.Structure):
    _fields_ = [
        ("fact_fk", C.c_void_p), ("fact_measure", C.c_void_p), ("n_fact", C.c_int64),
        ("dim", DimDesc), ("method", C.c_int32), ("flags", C.c_int32),
    ]


class JoinResult(C.Structure):
    _fields_ = [
        ("n_selected", C.c_int64), ("n_bloom_pass", C.c_int64), ("n_match", C.c_int64),
        ("sum_measure", C.c_int64), ("sum_product", C.c_uint64), ("time_build", C.c_float),
        ("time_probe", C.c_float), ("time_extra", C.c_float), ("time_total", C.c_float),
        ("stats", ProbeStats),
    ]


class StarDesc(C.Structure):
    _fields_ = [
        ("fact_fk", C.c_void_p * MAX_DIMS), ("fact_measure", C.c_void_p), ("n_fact", C.c_int64),
        ("n_dims", C.c_int32), ("method", C.c_int32), ("order", C.c_int32 * MAX_DIMS),
        ("bloom_order", C.c_int32 * MAX_DIMS), ("dims", DimDesc * MAX_DIMS),
        ("resident", C.c_void_p * MAX_DIMS), ("group_mult", C.c_int64 * MAX_DIMS),
        ("group_offset", C.c_int64), ("n_groups", C.c_int64), ("flags", C.c_int32),
        ("reserved0", C.c_int32), ("fact_measure_sub", C.c_void_p),
    ]


class Group(C.Structure):
    _fields_ = [
        ("g", C.c_int64), ("payload", C.c_int32 * MAX_DIMS), ("sum", C.c_int64),
        ("count", C.c_int64), ("sum_i32", C.c_int32), ("reserved0", C.c_int32),
    ]


class StarResult(C.Structure):
    _fields_ = [
        ("groups", C.POINTER(Group)), ("groups_cap", C.c_int64), ("n_groups", C.c_int64),
        ("n_rows", C.c_int64), ("sum_total", C.c_int64), ("time_query_ms", C.c_float),
        ("time_build_ms", C.c_float), ("time_probe_ms", C.c_float), ("status", C.c_int32),
        ("stats", ProbeStats),
    ]


class BatchOpts(C.Structure):
    _fields_ = [("use_graph", C.c_int32), ("probe_wg", C.c_int32), ("groups_cap", C.c_int64),
                ("launch_per_kind", C.c_int32), ("reserved0", C.c_int32), ("reserved", C.c_int64 * 1)]


class BloomChoice(C.Structure):
    _fields_ = [("bloom_hash", C.c_int32), ("bloom_layout", C.c_int32), ("bloom_k", C.c_int32),
                ("in_lds", C.c_int32), ("bloom_bits", C.c_int64), ("fpr", C.c_double), ("cost", C.c_double)]


class ServeCfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_int32), ("max_inflight", C.c_int32),
                ("groups_cap", C.c_int64), ("streams", C.c_int32), ("reserved0", C.c_int32),
                ("prewarm_groups", C.c_int64)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("launches", C.c_uint64), ("bytes_in_flight", C.c_uint64), ("peak_bytes_in_flight", C.c_uint64),
                ("max_batch_seen", C.c_uint64), ("reserved", C.c_uint64)]


class TuneModel(C.Structure):
    _fields_ = [("scan_l2", C.c_double), ("scan_big", C.c_double), ("probe_extra", C.c_double),
                ("line", C.c_double), ("gather", C.c_double), ("scan_mod", C.c_double),
                ("scan_blocked", C.c_double), ("gather_queue", C.c_double), ("gather_lds", C.c_double), ("gather_queue_lds", C.c_double),
                ("calibrated", C.c_int32),
                ("reserved", C.c_int32)]


STRUCTS = [DimDesc, DimInfo, ProbeStats, JoinDesc, JoinResult, StarDesc, Group, StarResult, BatchOpts,
           BloomChoice, ServeCfg, ServeStats, TuneModel]

# name -> (restype, argtypes); every entry point of include/lipjoin.h
SIGNATURES = {
    "lj_create": (C.c_int, [C.c_int32, C.c_uint64, C.c_void_p, C.POINTER(C.c_void_p)]),
    "lj_destroy": (None, [C.c_void_p]),
    "lj_last_error": (C.c_char_p, [C.c_void_p]),
    "lj_status_str": (C.c_char_p, [C.c_int]),
    "lj_abi_version": (C.c_int32, []),
    "lj_abi_struct_size": (C.c_int64, [C.c_int32]),
    "lj_sync": (C.c_int, [C.c_void_p]),
    "lj_bytes_in_use": (C.c_uint64, [C.c_void_p]),
    "lj_dim_build": (C.c_int, [C.c_void_p, C.POINTER(DimDesc), C.POINTER(C.c_void_p)]),
    "lj_dim_get_info": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(DimInfo)]),
    "lj_dim_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]),
    "lj_dim_release": (None, [C.c_void_p, C.c_void_p]),
    "lj_hash_join": (C.c_int, [C.c_void_p, C.POINTER(JoinDesc), C.POINTER(JoinResult)]),
    "lj_star_join": (C.c_int, [C.c_void_p, C.POINTER(StarDesc), C.POINTER(StarResult)]),
    "lj_star_join_noalloc": (C.c_int, [C.c_void_p, C.POINTER(StarDesc), C.POINTER(StarResult)]),
    "lj_batch_create": (C.c_int, [C.c_void_p, C.POINTER(StarDesc), C.c_int32, C.POINTER(BatchOpts),
                                  C.POINTER(C.c_void_p)]),
    "lj_batch_run": (C.c_int, [C.c_void_p, C.POINTER(StarResult)]),
    "lj_batch_destroy": (None, [C.c_void_p]),
    "lj_tune_bloom": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_int32, C.POINTER(BloomChoice)]),
    "lj_tune_calibrate": (C.c_int, [C.c_void_p]),
    "lj_tune_get_model": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lj_tune_set_model": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lj_serve_start": (C.c_int, [C.c_void_p, C.POINTER(ServeCfg), C.POINTER(C.c_void_p)]),
    "lj_serve_submit": (C.c_int, [C.c_void_p, C.POINTER(StarDesc), C.POINTER(C.c_uint64)]),
    "lj_serve_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(StarResult), C.c_int32]),
    "lj_serve_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServeStats)]),
    "lj_serve_stop": (C.c_int, [C.c_void_p]),
    "lj_gen_host": (C.c_int, [C.c_int32, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                              C.c_int64, C.c_void_p]),
    "lj_gen_device": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64,
                                C.c_int64, C.c_int64, C.c_void_p]),
    "lj_gen_shuffle_host": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, C.c_void_p]),
    "lj_gen_reference_host": (C.c_int, [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lj_gen_reference_star_host": (C.c_int, [C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path=None):
    """Load liblipjoin.so (once).  Raises LibraryMissing if it was not built:
    there is no fallback implementation."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibraryMissing(
            f"{p} is missing: build it with `make -C crystal-opt_amd` (or __graft_entry__.build())")
    # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7.  If
    # torch is installed it is imported first, so that the library binds to
    # torch's copy (same soname) instead of loading a second runtime, which
    # then finds no device once torch has opened it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check_layout(lib=None):
    """Compare the ctypes struct sizes with the library's sizeof()."""
    lib = lib or load()
    bad = []
    for i, st in enumerate(STRUCTS):
        n = lib.lj_abi_struct_size(i)
        if n != C.sizeof(st):
            bad.append((st.__name__, C.sizeof(st), n))
    return bad
