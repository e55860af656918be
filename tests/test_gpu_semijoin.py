"""Semi-joins through an exact filter (LJ_JOIN_SUM_ONLY, one-group star joins).

When a one-dimension LIP join needs no payload (SUM(measure) and counts only)
and its modulo filter has at least ht_len bits, the probe lets the filter
decide the table probe of every key inside the table's domain.  These tests
pin that shortcut to the reference semantics through the oracle on the cases
where the filter and the table could disagree: the key-0 sentinel (a selected
key 0 with payload 0 never matches, join.cuh:123-128), fact keys past the
table's domain (they wrap into the table, join.cuh:27), selected dimension
keys past the domain (the build counts them and the probe keeps the table
probe), filters of exactly ht_len bits and of a larger power of two, and both
the streamed-measure (pass rate above a quarter) and queued paths.  The
counts, sums and per-stage line statistics must equal the oracle's, and the
same join with the shortcut switched off (LJ_EXACT_FILTER=0).
"""
import os

import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O
from q33_helpers import to_device

pytestmark = pytest.mark.gpu

N_DIM = 100_000          # ht_len; 3125 words of 32 bits: a modulo filter of exactly ht_len bits


def _flush(msg):
    print(msg, flush=True)


def _case(seed, thr, wrap_dim=False):
    rng = np.random.default_rng(seed)
    keys = np.arange(N_DIM, dtype=np.int32)
    if wrap_dim:
        keys[-500:] = N_DIM + 7 + np.arange(500, dtype=np.int32) * 3   # selected keys past the domain
    attr = rng.integers(0, 100, N_DIM, dtype=np.int32)
    attr[0] = 0                                    # key 0 selected with payload 0: the sentinel
    n = (1 << 20) + 77
    fk = rng.integers(0, N_DIM, n, dtype=np.int32)
    fk[rng.integers(0, n, 3000)] = 0               # many rows on key 0
    far = rng.integers(0, n, 4000)
    fk[far] = rng.integers(N_DIM, 4 * N_DIM, far.size, dtype=np.int32)   # fact keys past the domain
    meas = rng.integers(1, 10001, n, dtype=np.int32)
    return keys, attr, fk, meas, thr


CASES = [
    # (selectivity threshold on attr < thr, filter bits, dimension keys past the domain)
    (60, N_DIM, False),          # streamed measure, filter of exactly ht_len bits
    (60, 1 << 17, False),        # streamed measure, power-of-two filter past ht_len
    (8, N_DIM, False),           # queued survivors
    (8, 1 << 17, False),
    (60, 1 << 17, True),         # aliased dimension keys: the table decides
]


@pytest.mark.parametrize("thr,bits,wrap", CASES, ids=[f"s{t}-b{b}-{'wrap' if w else 'in'}" for t, b, w in CASES])
def test_sum_only_hash_join_against_oracle(ctx, thr, bits, wrap):
    keys, attr, fk, meas, _ = _case(1234 + thr + bits, thr, wrap)
    _flush(f"semi-join hash join thr={thr} bits={bits} wrap={wrap}")
    pred = ((L.COMB_SET, L.CMP_LT, thr),)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    desc = L.dim_desc(dev["keys"], N_DIM, N_DIM, attr=dev["attr"], predicate=pred, bloom=("mod", bits))
    r = ctx.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="lip", sum_only=True)
    full = ctx.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="lip")
    spec = O.DimSpec(keys, attr, N_DIM, pred=list(pred), bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1))
    ref = O.single_join(spec, fk, meas, method=O.LIP)
    got = (r["n_selected"], r["n_bloom_pass"], r["n_match"], r["sum_measure"])
    assert got == (ref["n_selected"], ref["n_bloom_pass"], ref["n_match"], ref["sum_measure"])
    assert r["sum_product"] == 0
    if not wrap:   # an aliased slot's payload is its last writer's: order-dependent (join.cuh:276)
        assert full["sum_product"] == ref["sum_product"]
    assert (full["n_match"], full["sum_measure"]) == (ref["n_match"], ref["sum_measure"])
    groups, lines, rows, _ = O.star_join([spec], [fk], meas, (0,), method=O.LIP)
    assert r["stats"]["lines_in"][:2] == lines[:2]
    assert r["stats"]["rows_in"][:2] == rows[:2]


@pytest.mark.parametrize("thr", [60, 8])
def test_one_group_star_join_resident_against_oracle(ctx, thr):
    """A one-group star join over a resident dimension (the serving and batch
    shape): the shortcut is planned from the dimension's counters."""
    bits = 1 << 17
    keys, attr, fk, meas, _ = _case(99 + thr, thr)
    _flush(f"semi-join resident star join thr={thr}")
    pred = ((L.COMB_SET, L.CMP_LT, thr),)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    dim = ctx.build_dim(L.dim_desc(dev["keys"], N_DIM, N_DIM, attr=dev["attr"], predicate=pred, bloom=("mod", bits)))
    q = L.star_desc([dev["fk"]], dev["meas"], fk.size, None, (0,), method="lip", resident=[dim])
    res = ctx.star_join(q)
    dim.release()
    spec = O.DimSpec(keys, attr, N_DIM, pred=list(pred), bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1))
    groups, lines, rows, _ = O.star_join([spec], [fk], meas, (0,), method=O.LIP)
    assert (res["n_rows"], res["sum_total"]) == (groups[0][3], groups[0][2])
    assert res["stats"]["lines_in"][:2] == lines[:2]


def test_shortcut_off_gives_the_same(ctx):
    """The same joins with the shortcut switched off (a context created with
    LJ_EXACT_FILTER=0): identical results."""
    keys, attr, fk, meas, _ = _case(7, 60)
    pred = ((L.COMB_SET, L.CMP_LT, 60),)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    desc = L.dim_desc(dev["keys"], N_DIM, N_DIM, attr=dev["attr"], predicate=pred, bloom=("mod", N_DIM))
    on = ctx.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="lip", sum_only=True)
    os.environ["LJ_EXACT_FILTER"] = "0"
    try:
        with L.Context(0) as c2:
            off = c2.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="lip", sum_only=True)
    finally:
        del os.environ["LJ_EXACT_FILTER"]
    key = lambda r: (r["n_bloom_pass"], r["n_match"], r["sum_measure"], r["stats"]["lines_in"][:2])
    assert key(on) == key(off)
    torch.cuda.synchronize()


def test_tuner_semi_join_prefers_exact_filter(ctx):
    """With no gathers to save, the tuner's exact modulo filter (at least the
    dimension's rows in bits) wins at a high pass rate."""
    ch = ctx.tune_bloom(1 << 21, 1 << 22, 0.5, semi_join=True)
    assert ch["bloom_hash"] == L.BLOOM_MOD and ch["bloom_bits"] >= 1 << 22 and ch["fpr"] == 0.0


# ---- a filter too large for LDS, staged in part ------------------------------------
# A one-dimension semi-join whose exact modulo filter does not fit in LDS
# stages the filter's first words (what the LDS holds, about 1.2 Mbit) when
# the fact rows repay the copy (n_fact >= 4 x CUs x filter words / 32: 4.2 M
# rows for a 2^22-bit filter on 256 CUs); the probe reads the words past them
# from the global filter.  2^23 fact rows over a 2^22-key dimension take that
# path; a context with LJ_PARTIAL_PREFIX=0 reads the whole filter from global
# memory.  The joins with payloads (not sum-only) keep the whole filter in
# global memory either way: they check that the rule leaves them alone.
N_BIG = 1 << 22


def _big_case(seed, thr, wrap_dim=False):
    rng = np.random.default_rng(seed)
    keys = np.arange(N_BIG, dtype=np.int32)
    if wrap_dim:
        keys[-700:] = N_BIG + 11 + np.arange(700, dtype=np.int32) * 5
    attr = rng.integers(0, 100, N_BIG, dtype=np.int32)
    attr[0] = 0
    n = (1 << 23) + 19
    fk = rng.integers(0, N_BIG, n, dtype=np.int32)
    fk[rng.integers(0, n, 5000)] = 0
    far = rng.integers(0, n, 6000)
    fk[far] = rng.integers(N_BIG, 3 * N_BIG, far.size, dtype=np.int32)
    meas = rng.integers(1, 10001, n, dtype=np.int32)
    return keys, attr, fk, meas


BIG_CASES = [
    # (threshold, filter bits, sum only, dimension keys past the domain)
    (50, 1 << 22, True, False),      # exact filter, streamed measure
    (5, 1 << 22, True, False),       # exact filter, queued survivors
    (50, 1 << 22, True, True),       # aliased dimension keys: the table decides
    (50, 1 << 22, False, False),     # a filtered join with payloads: not staged
    (20, 3_000_017, False, False),   # a modulo filter of fewer bits than keys: not exact, not staged
    (20, (1 << 22) + 4099, True, False),  # exact, words not a power of two
]


@pytest.mark.parametrize("thr,bits,sum_only,wrap", BIG_CASES,
                         ids=[f"s{t}-b{b}-{'semi' if s else 'full'}-{'wrap' if w else 'in'}"
                              for t, b, s, w in BIG_CASES])
def test_partially_staged_filter_against_oracle(ctx, thr, bits, sum_only, wrap):
    keys, attr, fk, meas = _big_case(4321 + thr + bits % 1000, thr, wrap)
    _flush(f"partially staged filter thr={thr} bits={bits} sum_only={sum_only} wrap={wrap}")
    pred = ((L.COMB_SET, L.CMP_LT, thr),)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    desc = L.dim_desc(dev["keys"], N_BIG, N_BIG, attr=dev["attr"], predicate=pred, bloom=("mod", bits))
    r = ctx.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="lip", sum_only=sum_only)
    spec = O.DimSpec(keys, attr, N_BIG, pred=list(pred), bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1))
    ref = O.single_join(spec, fk, meas, method=O.LIP, threads=os.cpu_count() or 1)
    got = (r["n_selected"], r["n_bloom_pass"], r["n_match"], r["sum_measure"])
    assert got == (ref["n_selected"], ref["n_bloom_pass"], ref["n_match"], ref["sum_measure"])
    if not sum_only and not wrap:
        assert r["sum_product"] == ref["sum_product"]
    os.environ["LJ_PARTIAL_PREFIX"] = "0"
    try:
        with L.Context(0) as c2:
            whole = c2.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="lip", sum_only=sum_only)
    finally:
        del os.environ["LJ_PARTIAL_PREFIX"]
    key = lambda x: (x["n_bloom_pass"], x["n_match"], x["sum_measure"], x["sum_product"], x["stats"]["lines_in"][:2])
    assert key(r) == key(whole)
    torch.cuda.synchronize()
