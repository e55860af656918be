"""Pins the oracle's PHT_1 driver path and its textbook (open-addressed,
key-compare) table before they are trusted as the checker.

* PHT_1 (crystal/join.cuh:29-106,191-258): in a star join, a key-only table
  must keep exactly the rows a PHT_2 keeps when no key is 0 (the two differ
  only by the zero sentinel), with payload 0.
* Textbook table (not in the reference): its hits and payloads equal a
  Python dictionary of the selected keys (the last value of a repeated key
  wins), under collisions, a full table, one slot, and the excluded key.
* The documented divergence from the reference's perfect hash (Appendix A of
  the survey: no key compare, 0 = empty): key 0 with value 0 never hits in
  PHT_2 but hits in the textbook table; a fact key outside the dimension's
  key range aliases onto another key's slot in PHT_2 and misses in the
  textbook table.  On dense, nonzero, in-domain keys the two agree.
"""
import numpy as np
import pytest

import oracle as O

A0 = 0x9E3779B1


def _home(key, cap):
    lg = cap.bit_length() - 1
    return 0 if lg == 0 else ((key * A0) & 0xFFFFFFFF) >> (32 - lg)


@pytest.mark.parametrize("cap", [1, 2, 8, 1 << 10, 1 << 16])
def test_textbook_home_slot_closed_form(cap):
    for key in (0, 1, -1, 7, 123456789, -(1 << 31) + 1, (1 << 31) - 1):
        assert O.tb_home(key, cap) == _home(key & 0xFFFFFFFF, cap)


@pytest.mark.parametrize("n,cap,dups", [(3000, 1 << 13, False), (3000, 1 << 12, True), (1024, 1024, False),
                                        (1, 1, False), (50, 64, True)])
def test_textbook_equals_dictionary(n, cap, dups):
    rng = np.random.default_rng(n + cap)
    if dups:
        keys = rng.integers(-5000, 5000, n).astype(np.int32)       # repeats: last value wins
    else:
        keys = rng.choice(np.arange(-50000, 50000, dtype=np.int32), n, replace=False)
    vals = rng.integers(-1000, 1000, n).astype(np.int32)
    flags = (rng.random(n) < 0.8).astype(np.uint8)
    ht, oob = O.tb_build(keys, vals, flags, cap)
    ref = {}
    for k, v, f in zip(keys.tolist(), vals.tolist(), flags.tolist()):
        if f:
            ref[k] = v
    assert oob == 0 and len(ref) <= cap
    # every slot is empty or holds one selected key, each key once
    stored = (ht[0::2].view(np.uint32) ^ np.uint32(0x80000000)).view(np.int32)[ht[0::2] != 0]
    assert sorted(stored.tolist()) == sorted(ref)
    for k in set(keys.tolist()) | set(range(-60, 60)):
        hit, pay = O.tb_probe(ht, cap, k)
        assert hit == (k in ref) and (not hit or pay == ref[k]), k


def test_textbook_full_table_and_excluded_key():
    keys = np.array([5, 6, 7, 8, -(1 << 31)], dtype=np.int32)
    vals = np.arange(5, dtype=np.int32) + 1
    ht, oob = O.tb_build(keys, vals, None, 4)
    assert oob == 1                                  # INT32_MIN is outside the domain
    assert [O.tb_probe(ht, 4, k) for k in (5, 6, 7, 8)] == [(True, 1), (True, 2), (True, 3), (True, 4)]
    assert O.tb_probe(ht, 4, 9) == (False, None)     # full table: the probe ends after cap slots
    assert O.tb_probe(ht, 4, -(1 << 31)) == (False, None)
    ht2, oob2 = O.tb_build(np.array([1, 2, 3], dtype=np.int32), np.ones(3, np.int32), None, 2)
    assert oob2 == 1                                 # no room for the third key


def _dims(keys, vals, ht_len, kind, payload_key=False, keys_min=0):
    return O.DimSpec(keys, vals, ht_len, payload=keys if payload_key else None, keys_min=keys_min,
                     ht_kind=kind)


def test_pht1_keeps_the_pht2_rows_with_payload_zero():
    rng = np.random.default_rng(3)
    nd, nf = 4000, 30000
    keys = (rng.permutation(nd) + 1).astype(np.int32)       # no key 0
    attr = rng.integers(0, 10, nd).astype(np.int32)
    fk = rng.integers(1, nd + 1, nf).astype(np.int32)
    meas = rng.integers(1, 1000, nf).astype(np.int32)
    pred = [(O.COMB_SET, O.CMP_LT, 4)]
    for method in (O.NAIVE, O.LIP):
        d2 = O.DimSpec(keys, attr, nd, pred=pred)
        d1 = O.DimSpec(keys, attr, nd, pred=pred, ht_kind=O.HT_PHT1)
        g2, l2, r2, _ = O.star_join([d2], [fk], meas, (0,), method=method, n_groups=1)
        g1, l1, r1, _ = O.star_join([d1], [fk], meas, (0,), method=method, n_groups=1)
        assert [(g, s, c) for g, _, s, c in g1] == [(g, s, c) for g, _, s, c in g2]
        assert (l1, r1) == (l2, r2) and g1[0][1] == (0,)
        # and PHT_1 never hits key 0 (the sentinel), whatever its value
        k0 = np.array([0, 5], dtype=np.int32)
        s1 = O.single_join(O.DimSpec(k0, np.array([3, 3], np.int32), 8, ht_kind=O.HT_PHT1),
                           np.array([0, 0, 5], np.int32), np.array([1, 2, 4], np.int32), method=method)
        assert (s1["n_match"], s1["sum_measure"]) == (1, 4)


def test_textbook_agrees_with_compat_on_dense_in_domain_keys():
    rng = np.random.default_rng(4)
    nd, nf = 5000, 40000
    keys = (rng.permutation(nd) + 100).astype(np.int32)
    attr = rng.integers(0, 50, nd).astype(np.int32)
    fk = rng.integers(100, 100 + nd, nf).astype(np.int32)
    meas = rng.integers(1, 1000, nf).astype(np.int32)
    pred = [(O.COMB_SET, O.CMP_GTE, 20)]
    for method in (O.NAIVE, O.LIP):
        compat = O.DimSpec(keys, attr, nd, pred=pred, keys_min=100)
        tb = O.DimSpec(keys, attr, 1 << 14, pred=pred, ht_kind=O.HT_TEXTBOOK)
        gc, _, rc, _ = O.star_join([compat], [fk], meas, (0,), method=method, group_mult=(1, 0, 0, 0), n_groups=50)
        gt, _, rt, _ = O.star_join([tb], [fk], meas, (0,), method=method, group_mult=(1, 0, 0, 0), n_groups=50)
        assert gc == gt and rc == rt


def test_textbook_diverges_from_compat_on_zero_and_out_of_domain():
    """Key 0 with value 0: PHT_2's slot stays all zero, so it never hits
    (join.cuh:123-128); the textbook table stores it.  A fact key k + len
    lands on key k's slot in PHT_2 (no key compare) but misses in the
    textbook table."""
    keys = np.array([0, 1, 2, 3], dtype=np.int32)
    vals = np.array([0, 10, 20, 30], dtype=np.int32)
    fk = np.array([0, 0, 1, 5, 6, 3], dtype=np.int32)     # 5 -> slot 1, 6 -> slot 2 under (k % 4)
    meas = np.array([1, 2, 4, 8, 16, 32], dtype=np.int32)
    compat = O.single_join(O.DimSpec(keys, vals, 4), fk, meas, method=O.NAIVE)
    tb = O.single_join(O.DimSpec(keys, vals, 8, ht_kind=O.HT_TEXTBOOK), fk, meas, method=O.NAIVE)
    assert (compat["n_match"], compat["sum_measure"]) == (4, 4 + 8 + 16 + 32)
    assert (tb["n_match"], tb["sum_measure"]) == (4, 1 + 2 + 4 + 32)
    # the product sum uses the payload of each hit: 5 and 6 take the values of keys 1 and 2
    assert compat["sum_product"] == 4 * 10 + 8 * 10 + 16 * 20 + 32 * 30
    assert tb["sum_product"] == 1 * 0 + 2 * 0 + 4 * 10 + 32 * 30
