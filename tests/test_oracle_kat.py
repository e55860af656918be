"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each test states what fixes the expected answer: a known-answer test taken
from the reference's own test setup, a closed form, or a brute-force
recomputation by a different algorithm (nested loops over Python dicts) on
small inputs.  The reference ships no golden vectors for this path.
"""
import random

import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------------------
# Bloom filter, HASH1 (k = 1): the reference's only Bloom unit test,
# src/ops/bloom.cu:131-156 -- keys 0..1023, insert keys with i % 3 == 0 into a
# 16-word (512-bit) filter, then probe all 1024 keys.  The file prints TP/FN/
# TN/FP without asserting; the closed form is: bit b < 512 is set iff b or
# b + 512 is a multiple of 3, i.e. b % 3 in {0, 1} (512 % 3 == 2).
def test_bloom_mod_kat_from_reference_test():
    spec = O.bloom_spec(O.BLOOM_MOD, nbits=16 * 32, k=1)
    keys = np.arange(1024, dtype=np.int32)
    flags = (keys % 3 == 0).astype(np.uint8)
    words = O.bloom_build(spec, keys, flags)
    assert words.size == 16
    expect_bits = [b for b in range(512) if b % 3 in (0, 1)]
    assert len(expect_bits) == 342
    got_bits = [b for b in range(512) if (int(words[b >> 5]) >> (b & 31)) & 1]
    assert got_bits == expect_bits
    # the words repeat with period 3 (96 bits = 3 words)
    assert [hex(int(w)) for w in words[:3]] == ["0xdb6db6db", "0xb6db6db6", "0x6db6db6d"]
    assert all(words[i] == words[i % 3] for i in range(15))
    tp = fn = tn = fp = 0
    for k in range(1024):
        found = O.bloom_probe(spec, words, k)
        inserted = k % 3 == 0
        tp += inserted and found
        fn += inserted and not found
        tn += (not inserted) and not found
        fp += (not inserted) and found
    assert (tp, fn, tn, fp) == (342, 0, 340, 342)


def test_bloom_mod_keys_min_and_domain():
    # HASH1(X, Y, Z) = (X - Z) % Y; the date filter of q33 uses Z = 19920101
    spec = O.bloom_spec(O.BLOOM_MOD, nbits=65536, k=1, keys_min=19920101)
    assert O.bloom_bits(spec, 19920101) == [0]
    assert O.bloom_bits(spec, 19981230) == [19981230 - 19920101]
    assert O.bloom_bits(spec, 19920101 + 65536 + 7) == [7]
    assert O.bloom_bits(spec, 19920100) == []          # below keys_min: no bit (reference: UB)
    # bits = 33 rounds up to 2 words, as (bits + 31) / 32 in bloom_join.cu:128
    spec2 = O.bloom_spec(O.BLOOM_MOD, nbits=33, k=1)
    assert O.bloom_nwords(spec2) == 2
    assert O.bloom_bits(spec2, 64 + 5) == [5]


# ---------------------------------------------------------------------------
# Multiplicative family (not in the reference; defined by this project):
# h_i(x) = (x * A_i mod 2^32) >> (32 - log2 m).  Known answers computed by hand
# from the constants 0x9E3779B1, 0x85EBCA77, 0xC2B2AE3D, 0x27D4EB2F.
def test_bloom_mult_known_values():
    m = 1 << 16
    std3 = O.bloom_spec(O.BLOOM_MULT, nbits=m, k=3)
    assert O.bloom_bits(std3, 0) == [0, 0, 0]
    assert O.bloom_bits(std3, 1) == [0x9E37, 0x85EB, 0xC2B2]       # top 16 bits of A_i
    assert O.bloom_bits(std3, 2) == [0x3C6E, 0x0BD7, 0x8565]       # 2*A mod 2^32 = 0x3C6EF362, 0x0BD794EE, 0x85655C7A
    # keys_min shifts the argument
    std1 = O.bloom_spec(O.BLOOM_MULT, nbits=m, k=1, keys_min=100)
    assert O.bloom_bits(std1, 101) == [0x9E37]
    # blocked: word = (x*A0) >> (32 - log2(m/32)) = top 11 bits; bit_i = top 5 bits of x*A_{i+1}
    blk = O.bloom_spec(O.BLOOM_MULT, nbits=m, k=3, layout=O.LAYOUT_BLOCKED)
    w = 0x9E3779B1 >> 21
    assert O.bloom_bits(blk, 1) == [w * 32 + (0x85EBCA77 >> 27), w * 32 + (0xC2B2AE3D >> 27),
                                    w * 32 + (0x27D4EB2F >> 27)]


@pytest.mark.parametrize("layout", [O.LAYOUT_STANDARD, O.LAYOUT_BLOCKED])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_bloom_mult_invariants(k, layout):
    rng = np.random.default_rng(7 + k)
    m = 1 << 12
    spec = O.bloom_spec(O.BLOOM_MULT, nbits=m, k=k, layout=layout)
    keys = rng.integers(-2**31, 2**31 - 1, size=300, dtype=np.int64).astype(np.int32)
    words = O.bloom_build(spec, keys)
    # no false negatives
    assert all(O.bloom_probe(spec, words, int(x)) for x in keys)
    # at most k bits per key
    assert sum(bin(int(w)).count("1") for w in words) <= k * keys.size
    # independent of insertion order
    perm = rng.permutation(keys.size)
    assert np.array_equal(words, O.bloom_build(spec, keys[perm]))
    # blocked layout keeps each key's bits inside one word
    if layout == O.LAYOUT_BLOCKED:
        for x in keys[:50]:
            assert len({b >> 5 for b in O.bloom_bits(spec, int(x))}) == 1


# ---------------------------------------------------------------------------
def test_round_up_pow2_closed_form():
    # ssb_utils.h:85-96
    assert [O.round_up_pow2(n) for n in (0, 1, 2, 3, 4, 5, 40000, 61130, 600000, 2**31)] == \
        [1, 1, 2, 4, 4, 8, 65536, 65536, 1048576, 2**31]


def test_hash_pht_truncating_remainder():
    # HASH(X,Y,Z) ((X-Z) % Y), join.cuh:27
    assert O.hash_pht(600000, 600000) == 0
    assert O.hash_pht(599999, 600000) == 599999
    assert O.hash_pht(19981230, 61130, 19920101) == 61129
    assert O.hash_pht(5, 3, 0) == 2
    assert O.hash_pht(-1, 10, 0) == -1          # negative C remainder -> out of bounds in the reference
    assert O.hash_pht(19920100, 61130, 19920101) == -1


def test_predicate_semantics():
    attr = np.arange(1985, 2001, dtype=np.int32)
    # BlockPredGTE then BlockPredLTE: the second call overwrites (q33/0-lip.cu:192-195)
    f = O.predicate(attr, [(O.COMB_SET, O.CMP_GTE, 1992), (O.COMB_SET, O.CMP_LTE, 1997)])
    assert [int(a) for a, x in zip(attr, f) if x] == list(range(1985, 1998))
    f = O.predicate(attr, [(O.COMB_SET, O.CMP_GTE, 1992), (O.COMB_AND, O.CMP_LTE, 1997)])
    assert [int(a) for a, x in zip(attr, f) if x] == list(range(1992, 1998))
    f = O.predicate(attr, [(O.COMB_SET, O.CMP_EQ, 1990), (O.COMB_OR, O.CMP_EQ, 1995)])
    assert [int(a) for a, x in zip(attr, f) if x] == [1990, 1995]
    for cmp, fn in [(O.CMP_LT, lambda a: a < 1993), (O.CMP_GT, lambda a: a > 1993),
                    (O.CMP_LTE, lambda a: a <= 1993), (O.CMP_GTE, lambda a: a >= 1993),
                    (O.CMP_EQ, lambda a: a == 1993)]:
        f = O.predicate(attr, [(O.COMB_SET, cmp, 1993)])
        assert list(f.astype(bool)) == [fn(int(a)) for a in attr]
    assert O.predicate(attr, []).all()   # InitFlags


# ---------------------------------------------------------------------------
def test_pht2_dense_keys_equal_direct_index_join():
    """On dense in-domain keys PHT_2 equals a textbook direct-index join."""
    rng = np.random.default_rng(3)
    n = 5000
    keys = rng.permutation(np.arange(1, n + 1)).astype(np.int32)
    vals = rng.integers(1, 250, n).astype(np.int32)
    flags = (rng.random(n) < 0.3).astype(np.uint8)
    ht, oob = O.pht2_build(keys, vals, flags, n, 0)
    assert oob == 0
    direct = {int(k): int(v) for k, v, f in zip(keys, vals, flags) if f}
    for key in range(0, n + 2):
        hit, pay = O.pht2_probe(ht, n, 0, key)
        if 1 <= key <= n:
            assert (hit, pay) == ((True, direct[key]) if key in direct else (False, None))
    # layout: pair (key, val) at slot key % n
    k0 = int(keys[np.nonzero(flags)[0][0]])
    assert ht[2 * (k0 % n)] == k0 and ht[2 * (k0 % n) + 1] == direct[k0]


def test_pht_zero_sentinel_quirk():
    # join.cuh:123-128: a (key 0, val 0) pair is indistinguishable from an empty slot
    ht, _ = O.pht2_build(np.array([0, 1], np.int32), np.array([0, 0], np.int32), None, 4)
    assert O.pht2_probe(ht, 4, 0, 0) == (False, None)
    assert O.pht2_probe(ht, 4, 0, 1) == (True, 0)          # key 1 is nonzero: hit with payload 0
    ht, _ = O.pht2_build(np.array([0], np.int32), np.array([5], np.int32), None, 4)
    assert O.pht2_probe(ht, 4, 0, 0) == (True, 5)
    assert O.pht2_probe(ht, 4, 0, 4) == (True, 5)          # no key compare: 4 aliases slot 0
    # PHT_1: key 0 never matches
    ht1, _ = O.pht1_build(np.array([0, 3], np.int32), None, 4)
    assert not O.pht1_probe(ht1, 4, 0, 0) and O.pht1_probe(ht1, 4, 0, 3)


def test_pht2_last_writer_wins_on_shared_slot():
    # join.cuh:275-276: CAS keeps the first key, the value store is unconditional
    ht, _ = O.pht2_build(np.array([3, 7], np.int32), np.array([10, 20], np.int32), None, 4)
    assert ht[2 * 3] == 3 and ht[2 * 3 + 1] == 20


# ---------------------------------------------------------------------------
def _brute_single(keys, attr, thr, fk, meas, bloom=None, bloom_words=None):
    sel = {int(k): int(a) for k, a in zip(keys, attr) if a < thr}
    n_match = s = prod = npass = 0
    for f, m in zip(fk, meas):
        f, m = int(f), int(m)
        if bloom is not None and not O.bloom_probe(bloom, bloom_words, f):
            continue
        npass += 1
        if f in sel and not (f == 0 and sel[f] == 0):
            n_match += 1
            s += m
            p = (m * sel[f]) & 0xFFFFFFFF
            p = p - (1 << 32) if p >= 1 << 31 else p
            prod = (prod + p) & 0xFFFFFFFFFFFFFFFF
    return n_match, s, prod, npass


@pytest.mark.parametrize("method", [O.NAIVE, O.LIP])
def test_single_join_brute_force(method):
    rng = np.random.default_rng(11)
    nd, nf = 3000, 20000
    keys = rng.permutation(nd).astype(np.int32)
    attr = rng.integers(0, 100, nd).astype(np.int32)
    fk = rng.integers(0, nd, nf).astype(np.int32)
    meas = rng.integers(1, 10001, nf).astype(np.int32)
    bloom = O.bloom_spec(O.BLOOM_MULT, nbits=1 << 10, k=2)
    dim = O.DimSpec(keys, attr, nd, pred=[(O.COMB_SET, O.CMP_LT, 10)], bloom=bloom)
    r = O.single_join(dim, fk, meas, method=method, threads=3)
    flags = O.predicate(attr, [(O.COMB_SET, O.CMP_LT, 10)])
    words = O.bloom_build(bloom, keys, flags)
    n_match, s, prod, npass = _brute_single(keys, attr, 10, fk, meas,
                                            bloom if method == O.LIP else None, words)
    assert r["n_selected"] == int(flags.sum())
    assert (r["n_match"], r["sum_measure"], r["sum_product"]) == (n_match, s, prod)
    assert r["n_bloom_pass"] == npass
    assert r["lines_fk"] == (nf + 31) // 32


def test_single_join_closed_forms():
    n = 4096
    keys = np.arange(1, n + 1, dtype=np.int32)
    attr = np.full(n, 7, np.int32)
    fk = np.tile(keys, 3)
    meas = np.arange(fk.size, dtype=np.int32)
    # every dimension row selected: every fact row matches
    r = O.single_join(O.DimSpec(keys, attr, n), fk, meas, method=O.NAIVE)
    assert r["n_match"] == fk.size and r["sum_measure"] == fk.size * (fk.size - 1) // 2
    assert r["sum_product"] == 7 * fk.size * (fk.size - 1) // 2
    # nothing selected: nothing matches
    r = O.single_join(O.DimSpec(keys, attr, n, pred=[(O.COMB_SET, O.CMP_LT, 0)]), fk, meas, method=O.NAIVE)
    assert (r["n_selected"], r["n_match"], r["sum_measure"], r["lines_measure"]) == (0, 0, 0, 0)


# ---------------------------------------------------------------------------
def _brute_star(dim_rows, dim_preds, fks, meas, gmult, goff, ngroups):
    """q33.sql semantics by nested loops: for each fact row, find the matching
    selected dimension row of every dimension by key equality."""
    tables = []
    for (keys, attr), pf in zip(dim_rows, dim_preds):
        tables.append({int(k): int(a) for k, a in zip(keys, attr) if pf(int(a))})
    out = {}
    for i in range(len(meas)):
        pays = []
        for t, f in zip(tables, fks):
            k = int(f[i])
            if k not in t:
                break
            pays.append(t[k])
        else:
            g = (goff + sum(p * m for p, m in zip(pays, gmult))) % ngroups
            s, c, _ = out.get(g, (0, 0, None))
            out[g] = (s + int(meas[i]), c + 1, tuple(pays))
    return sorted((g, p, s, c) for g, (s, c, p) in out.items())


def _small_star(seed, n_fact):
    rng = np.random.default_rng(seed)
    ns, nc, nd = 50, 300, 40
    s_keys = np.arange(1, ns + 1, dtype=np.int32)
    s_city = rng.integers(0, 10, ns).astype(np.int32)
    c_keys = np.arange(1, nc + 1, dtype=np.int32)
    c_city = rng.integers(0, 10, nc).astype(np.int32)
    d_keys = (1000 + np.arange(nd)).astype(np.int32)
    d_year = (1992 + np.arange(nd) // 6).astype(np.int32)
    fk_s = rng.integers(1, ns + 1, n_fact).astype(np.int32)
    fk_c = rng.integers(1, nc + 1, n_fact).astype(np.int32)
    fk_d = rng.choice(d_keys, n_fact).astype(np.int32)
    meas = rng.integers(1, 10**6, n_fact).astype(np.int32)
    return (s_keys, s_city, c_keys, c_city, d_keys, d_year), (fk_s, fk_c, fk_d, meas)


@pytest.mark.parametrize("method", [O.NAIVE, O.LIP])
@pytest.mark.parametrize("order", [(0, 1, 2), (0, 2, 1), (2, 0, 1), (2, 1, 0), (1, 2, 0), (1, 0, 2)])
def test_star_join_brute_force(method, order):
    (s_keys, s_city, c_keys, c_city, d_keys, d_year), (fk_s, fk_c, fk_d, meas) = _small_star(5, 3000)
    city = [(O.COMB_SET, O.CMP_EQ, 3), (O.COMB_OR, O.CMP_EQ, 7)]
    year = [(O.COMB_SET, O.CMP_GTE, 1992), (O.COMB_SET, O.CMP_LTE, 1996)]
    dims = [
        O.DimSpec(s_keys, s_city, 50, pred=city, bloom=O.bloom_spec(O.BLOOM_MOD, 64)),
        O.DimSpec(c_keys, c_city, 300, pred=city, bloom=O.bloom_spec(O.BLOOM_MOD, 512)),
        O.DimSpec(d_keys, d_year, 40, pred=year, keys_min=1000, bloom=O.bloom_spec(O.BLOOM_MOD, 64, keys_min=1000)),
    ]
    gm, go, ng = (10 * 7, 7, 1), -1992, 7 * 10 * 10
    groups, lines, rows, oob = O.star_join(dims, [fk_s, fk_c, fk_d], meas, order, method=method,
                                           group_mult=gm, group_offset=go, n_groups=ng, threads=2)
    expect = _brute_star([(s_keys, s_city), (c_keys, c_city), (d_keys, d_year)],
                         [lambda a: a in (3, 7), lambda a: a in (3, 7), lambda a: a <= 1996],
                         [fk_s, fk_c, fk_d], meas, gm, go, ng)
    assert groups == expect and oob == 0
    assert len(groups) > 3


def test_star_line_and_row_statistics_brute_force():
    (s_keys, s_city, c_keys, c_city, d_keys, d_year), (fk_s, fk_c, fk_d, meas) = _small_star(9, 5000)
    city = [(O.COMB_SET, O.CMP_LT, 4)]
    dims = [O.DimSpec(s_keys, s_city, 50, pred=city, bloom=O.bloom_spec(O.BLOOM_MULT, 1 << 8, 2)),
            O.DimSpec(c_keys, c_city, 300, pred=city, bloom=O.bloom_spec(O.BLOOM_MULT, 1 << 10, 1)),
            O.DimSpec(d_keys, d_year, 40, keys_min=1000)]
    for method in (O.NAIVE, O.LIP):
        order = (1, 2, 0)
        _, lines, rows, _ = O.star_join(dims, [fk_s, fk_c, fk_d], meas, order, method=method, n_groups=1)
        # recount: alive sets per stage, then distinct (row >> 5)
        sel = [set(int(k) for k, a in zip(s_keys, s_city) if a < 4),
               set(int(k) for k, a in zip(c_keys, c_city) if a < 4), set(int(k) for k in d_keys)]
        words = [O.bloom_build(dims[i].bloom, dims[i].keys, O.predicate(dims[i].attr, dims[i].pred))
                 if i < 2 else None for i in range(3)]
        fks = [fk_s, fk_c, fk_d]
        alive = np.ones(meas.size, bool)
        exp_rows, exp_lines = [], []
        for d in order:
            exp_rows.append(int(alive.sum()))
            exp_lines.append(len(set((np.nonzero(alive)[0] >> 5).tolist())))
            if method == O.NAIVE:
                ok = np.array([int(k) in sel[d] for k in fks[d]])
            else:
                ok = np.array([words[d] is None or O.bloom_probe(dims[d].bloom, words[d], int(k)) for k in fks[d]])
            alive &= ok
        if method == O.LIP:
            bloom_pass = int(alive.sum())
            for d in order:
                alive &= np.array([int(k) in sel[d] for k in fks[d]])
        exp_rows.append(int(alive.sum()))
        exp_lines.append(len(set((np.nonzero(alive)[0] >> 5).tolist())))
        assert rows[:4] == exp_rows and lines == exp_lines
        if method == O.LIP:
            assert rows[4] == bloom_pass


def test_oracle_threads_do_not_change_results():
    (s_keys, s_city, c_keys, c_city, d_keys, d_year), (fk_s, fk_c, fk_d, meas) = _small_star(2, 7777)
    dims = [O.DimSpec(s_keys, s_city, 50, pred=[(O.COMB_SET, O.CMP_LT, 5)]),
            O.DimSpec(c_keys, c_city, 300, pred=[(O.COMB_SET, O.CMP_LT, 5)]),
            O.DimSpec(d_keys, d_year, 40, keys_min=1000)]
    ref = O.star_join(dims, [fk_s, fk_c, fk_d], meas, (0, 1, 2), method=O.NAIVE, group_mult=(70, 7, 1),
                      group_offset=-1992, n_groups=700, threads=1)
    for t in (2, 3, 8):
        assert O.star_join(dims, [fk_s, fk_c, fk_d], meas, (0, 1, 2), method=O.NAIVE, group_mult=(70, 7, 1),
                           group_offset=-1992, n_groups=700, threads=t) == ref
    random.seed(0)
