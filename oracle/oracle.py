"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The shipped library never
uses it.  See oracle.h for what each function restates and how it is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
MAX_DIMS = 4

COMB_SET, COMB_AND, COMB_OR = 0, 1, 2
CMP_LT, CMP_GT, CMP_LTE, CMP_GTE, CMP_EQ = 0, 1, 2, 3, 4
BLOOM_NONE, BLOOM_MOD, BLOOM_MULT = 0, 1, 2
LAYOUT_STANDARD, LAYOUT_BLOCKED = 0, 1
NAIVE, LIP = 0, 1


class PredStep(C.Structure):
    _fields_ = [("comb", C.c_int32), ("cmp", C.c_int32), ("value", C.c_int32)]


class Bloom(C.Structure):
    _fields_ = [("hash", C.c_int32), ("layout", C.c_int32), ("k", C.c_int32),
                ("keys_min", C.c_int32), ("nbits", C.c_int64)]


HT_PHT2, HT_PHT1, HT_TEXTBOOK = 0, 1, 2


class Dim(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("attr", C.c_void_p), ("payload", C.c_void_p),
                ("n", C.c_int64), ("pred", C.POINTER(PredStep)), ("n_pred", C.c_int32),
                ("keys_min", C.c_int32), ("ht_len", C.c_int64), ("bloom", Bloom),
                ("ht_kind", C.c_int32), ("reserved0", C.c_int32)]


class JoinResult(C.Structure):
    _fields_ = [("n_selected", C.c_int64), ("n_bloom_pass", C.c_int64), ("n_match", C.c_int64),
                ("sum_measure", C.c_int64), ("sum_product", C.c_uint64), ("n_oob", C.c_int64),
                ("lines_fk", C.c_int64), ("lines_measure", C.c_int64)]


class Fact(C.Structure):
    _fields_ = [("fk", C.c_void_p * MAX_DIMS), ("measure", C.c_void_p), ("n", C.c_int64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        i32p, i64p, u32p, u8p = (C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint8))
        lib.orc_round_up_pow2.restype = C.c_uint32
        lib.orc_round_up_pow2.argtypes = [C.c_uint32]
        lib.orc_hash_pht.restype = C.c_int64
        lib.orc_hash_pht.argtypes = [C.c_int32, C.c_int64, C.c_int32]
        lib.orc_predicate.argtypes = [C.c_void_p, C.c_int64, C.POINTER(PredStep), C.c_int32, C.c_void_p]
        lib.orc_pht1_build.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, i64p]
        lib.orc_pht1_probe.restype = C.c_int
        lib.orc_pht1_probe.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32]
        lib.orc_pht2_build.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                       C.c_int32, i64p]
        lib.orc_pht2_probe.restype = C.c_int
        lib.orc_pht2_probe.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, i32p]
        lib.orc_bloom_nwords.restype = C.c_int64
        lib.orc_bloom_nwords.argtypes = [C.POINTER(Bloom)]
        lib.orc_bloom_build.argtypes = [C.POINTER(Bloom), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, i64p]
        lib.orc_bloom_probe.restype = C.c_int
        lib.orc_bloom_probe.argtypes = [C.POINTER(Bloom), C.c_void_p, C.c_int32]
        lib.orc_bloom_bits.restype = C.c_int
        lib.orc_bloom_bits.argtypes = [C.POINTER(Bloom), C.c_int32, i64p]
        lib.orc_single_join.restype = C.c_int
        lib.orc_single_join.argtypes = [C.POINTER(Dim), C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                        C.POINTER(JoinResult)]
        lib.orc_tb_home.restype = C.c_int64
        lib.orc_tb_home.argtypes = [C.c_int32, C.c_int64]
        lib.orc_tb_build.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, i64p]
        lib.orc_tb_probe.restype = C.c_int
        lib.orc_tb_probe.argtypes = [C.c_void_p, C.c_int64, C.c_int32, i32p]
        lib.orc_star_join.restype = C.c_int64
        lib.orc_star_join.argtypes = [C.POINTER(Dim), C.c_int32, C.POINTER(Fact), i32p, i32p, C.c_int32, i64p,
                                      C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, i64p, i64p, i64p,
                                      C.c_int32]
        _ = (u32p, u8p)
        _lib = lib
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def pred_array(steps):
    arr = (PredStep * max(1, len(steps)))()
    for i, (comb, cmp, val) in enumerate(steps):
        arr[i].comb, arr[i].cmp, arr[i].value = comb, cmp, val
    return arr


def predicate(attr, steps):
    attr = _i32(attr)
    flags = np.zeros(attr.size, dtype=np.uint8)
    load().orc_predicate(attr.ctypes.data, attr.size, pred_array(steps), len(steps), flags.ctypes.data)
    return flags


def hash_pht(key, ht_len, keys_min=0):
    return load().orc_hash_pht(int(key), int(ht_len), int(keys_min))


def round_up_pow2(n):
    return load().orc_round_up_pow2(int(n))


def bloom_spec(hash_kind=BLOOM_NONE, nbits=0, k=1, layout=LAYOUT_STANDARD, keys_min=0):
    b = Bloom()
    b.hash, b.nbits, b.k, b.layout, b.keys_min = hash_kind, int(nbits), int(k), int(layout), int(keys_min)
    return b


def bloom_nwords(spec):
    return load().orc_bloom_nwords(C.byref(spec))


def bloom_build(spec, keys, flags=None):
    keys = _i32(keys)
    words = np.zeros(bloom_nwords(spec), dtype=np.uint32)
    oob = C.c_int64(0)
    f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    load().orc_bloom_build(C.byref(spec), keys.ctypes.data, None if f is None else f.ctypes.data, keys.size,
                           words.ctypes.data, C.byref(oob))
    return words


def bloom_probe(spec, words, key):
    return bool(load().orc_bloom_probe(C.byref(spec), words.ctypes.data, int(key)))


def bloom_bits(spec, key):
    out = (C.c_int64 * 4)()
    k = load().orc_bloom_bits(C.byref(spec), int(key), out)
    return [out[i] for i in range(k)]


def pht2_build(keys, vals, flags, ht_len, keys_min=0):
    keys, vals = _i32(keys), _i32(vals)
    ht = np.zeros(2 * int(ht_len), dtype=np.int32)
    oob = C.c_int64(0)
    f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    load().orc_pht2_build(keys.ctypes.data, vals.ctypes.data, None if f is None else f.ctypes.data, keys.size,
                          ht.ctypes.data, int(ht_len), int(keys_min), C.byref(oob))
    return ht, oob.value


def pht2_probe(ht, ht_len, keys_min, key):
    pay = C.c_int32(0)
    hit = load().orc_pht2_probe(ht.ctypes.data, int(ht_len), int(keys_min), int(key), C.byref(pay))
    return (True, pay.value) if hit else (False, None)


def pht1_build(keys, flags, ht_len, keys_min=0):
    keys = _i32(keys)
    ht = np.zeros(int(ht_len), dtype=np.int32)
    oob = C.c_int64(0)
    f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    load().orc_pht1_build(keys.ctypes.data, None if f is None else f.ctypes.data, keys.size, ht.ctypes.data,
                          int(ht_len), int(keys_min), C.byref(oob))
    return ht, oob.value


def pht1_probe(ht, ht_len, keys_min, key):
    return bool(load().orc_pht1_probe(ht.ctypes.data, int(ht_len), int(keys_min), int(key)))


def tb_home(key, cap):
    return load().orc_tb_home(int(key), int(cap))


def tb_build(keys, vals, flags, cap):
    keys, vals = _i32(keys), _i32(vals)
    ht = np.zeros(2 * int(cap), dtype=np.int32)
    oob = C.c_int64(0)
    f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    load().orc_tb_build(keys.ctypes.data, vals.ctypes.data, None if f is None else f.ctypes.data, keys.size,
                        ht.ctypes.data, int(cap), C.byref(oob))
    return ht, oob.value


def tb_probe(ht, cap, key):
    pay = C.c_int32(0)
    hit = load().orc_tb_probe(ht.ctypes.data, int(cap), int(key), C.byref(pay))
    return (True, pay.value) if hit else (False, None)


class DimSpec:
    """Host-side dimension: numpy columns + predicate + table/filter parameters."""

    def __init__(self, keys, attr, ht_len, pred=(), payload=None, keys_min=0, bloom=None, ht_kind=HT_PHT2):
        self.keys = _i32(keys)
        self.attr = _i32(attr) if attr is not None else None
        self.payload = _i32(payload) if payload is not None else None
        self.ht_len, self.pred, self.keys_min = int(ht_len), list(pred), int(keys_min)
        self.bloom = bloom if bloom is not None else bloom_spec()
        self.ht_kind = int(ht_kind)
        self._pred_arr = pred_array(self.pred)

    def c(self):
        d = Dim()
        d.keys = self.keys.ctypes.data
        d.attr = self.attr.ctypes.data if self.attr is not None else None
        d.payload = self.payload.ctypes.data if self.payload is not None else None
        d.n = self.keys.size
        d.pred = C.cast(self._pred_arr, C.POINTER(PredStep))
        d.n_pred = len(self.pred)
        d.keys_min, d.ht_len, d.bloom = self.keys_min, self.ht_len, self.bloom
        d.ht_kind = self.ht_kind
        return d


def single_join(dim, fk, measure, method=LIP, threads=1):
    fk, measure = _i32(fk), _i32(measure)
    r = JoinResult()
    d = dim.c()
    rc = load().orc_single_join(C.byref(d), fk.ctypes.data, measure.ctypes.data, fk.size, method, threads,
                                C.byref(r))
    if rc != 0:
        raise RuntimeError("orc_single_join failed")
    return {k: getattr(r, k) for k, _ in JoinResult._fields_}


def star_join(dims, fks, measure, order, method=LIP, bloom_order=None, group_mult=(0, 0, 0, 0),
              group_offset=0, n_groups=1, threads=1):
    """Returns (groups, lines, rows, n_oob); groups = [(g, payloads, sum, count)] ascending g."""
    nd = len(dims)
    fks = [_i32(f) for f in fks]
    measure = _i32(measure)
    fact = Fact()
    for i in range(nd):
        fact.fk[i] = fks[i].ctypes.data
    fact.measure, fact.n = measure.ctypes.data, measure.size
    cdims = (Dim * nd)(*[d.c() for d in dims])
    o = (C.c_int32 * MAX_DIMS)(*order)
    bo = (C.c_int32 * MAX_DIMS)(*(bloom_order if bloom_order is not None else order))
    gm = (C.c_int64 * MAX_DIMS)(*list(group_mult)[:MAX_DIMS])
    gsum = np.zeros(n_groups, dtype=np.int64)
    gcount = np.zeros(n_groups, dtype=np.int64)
    gpay = np.zeros(n_groups * MAX_DIMS, dtype=np.int32)
    lines = (C.c_int64 * (MAX_DIMS + 1))()
    rows = (C.c_int64 * (MAX_DIMS + 2))()
    oob = C.c_int64(0)
    hit = load().orc_star_join(cdims, nd, C.byref(fact), o, bo, method, gm, int(group_offset), int(n_groups),
                               gsum.ctypes.data, gcount.ctypes.data, gpay.ctypes.data, lines, rows, C.byref(oob),
                               threads)
    if hit < 0:
        raise RuntimeError("orc_star_join failed")
    nz = np.nonzero(gcount)[0]
    groups = [(int(g), tuple(int(x) for x in gpay[g * MAX_DIMS:g * MAX_DIMS + nd]), int(gsum[g]), int(gcount[g]))
              for g in nz]
    return groups, list(lines)[:nd + 1], list(rows)[:nd + 2], oob.value
