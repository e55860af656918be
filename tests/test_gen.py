"""Seeded generators (lj_gen_host): determinism, chunking, ranges, calendar."""
import datetime

import numpy as np

import lipjoin as L
from lipjoin import workloads as W


def test_counter_based_chunks_match_whole_column():
    whole = L.gen_host(L.GEN_UNIFORM, 123, 4, -50, 50, 100000)
    parts = np.concatenate([L.gen_host(L.GEN_UNIFORM, 123, 4, -50, 50, 30000, first=0),
                            L.gen_host(L.GEN_UNIFORM, 123, 4, -50, 50, 70000, first=30000)])
    assert np.array_equal(whole, parts)
    assert whole.min() == -50 and whole.max() == 50
    assert not np.array_equal(whole, L.gen_host(L.GEN_UNIFORM, 124, 4, -50, 50, 100000))
    assert not np.array_equal(whole, L.gen_host(L.GEN_UNIFORM, 123, 5, -50, 50, 100000))


def test_uniform_is_roughly_uniform():
    x = L.gen_host(L.GEN_UNIFORM, 9, 1, 0, 9, 200000)
    counts = np.bincount(x, minlength=10)
    assert counts.min() > 19000 and counts.max() < 21000
    full = L.gen_host(L.GEN_UNIFORM, 9, 2, -2**31, 2**31 - 1, 1000)   # range 2^32
    assert full.min() < -2**30 and full.max() > 2**30


def test_calendar_matches_python_datetime():
    keys = L.gen_host(L.GEN_DATEKEY_SEQ, 0, 0, 0, 0, W.Q33_D_LEN)
    years = L.gen_host(L.GEN_YEAR_OF_DAY, 0, 0, 0, 0, W.Q33_D_LEN)
    d0 = datetime.date(1992, 1, 1)
    exp = [(d0 + datetime.timedelta(days=i)) for i in range(W.Q33_D_LEN)]
    assert keys.tolist() == [int(d.strftime("%Y%m%d")) for d in exp]
    assert years.tolist() == [d.year for d in exp]
    assert keys[0] == 19920101 and keys[-1] == 19981230        # the SSB date table, 2556 days
    od = L.gen_host(L.GEN_DATEKEY, 7, 12, 0, W.Q33_D_LEN, 50000)
    assert set(od.tolist()) <= set(keys.tolist())
    assert len(set(od.tolist())) == W.Q33_D_LEN


def test_shuffle_is_a_permutation():
    a = np.arange(100000, dtype=np.int32)
    L.shuffle_host(a, 5, 0)
    assert np.array_equal(np.sort(a), np.arange(100000))
    assert (a != np.arange(100000)).mean() > 0.99
    b = np.arange(100000, dtype=np.int32)
    L.shuffle_host(b, 5, 0)
    assert np.array_equal(a, b)


def test_q33_fact_ranges():
    f = W.q33_host_fact(W.seed_of(3), 20000)
    assert f["lo_custkey"].min() >= 1 and f["lo_custkey"].max() <= 600000
    assert f["lo_suppkey"].min() >= 1 and f["lo_suppkey"].max() <= 40000
    assert f["lo_revenue"].min() >= 1 and f["lo_revenue"].max() <= 10**6
    assert f["lo_orderdate"].min() >= 19920101 and f["lo_orderdate"].max() <= 19981230
