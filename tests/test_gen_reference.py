"""The reference's single-join generator (src/ops/utils/generator.h:37-136),
as lj_gen_reference_host restates it, against an independent Python
restatement over the same C library's rand() sequence (called through
ctypes): srand(0), create_relation_pk (keys 1..n shuffled by the Knuth
shuffle with j = RAND_RANGE(i), generator.h:32,51-60; values 0..n-1), then
create_relation_fk (successive shuffled blocks of 1..n_dim and a shuffled
1..(n_fact % n_dim) tail; values 0..n_fact-1).  No GPU."""
import ctypes
import ctypes.util

import numpy as np
import pytest

import lipjoin as L

LIBC = ctypes.CDLL(ctypes.util.find_library("c"))
LIBC.rand.restype = ctypes.c_int
RAND_MAX = 2147483647          # glibc


def _rand_range(n):
    return int(LIBC.rand() / (RAND_MAX + 1.0) * n)


def _unique_shuffled(n):
    a = list(range(1, n + 1))
    for i in range(n - 1, 0, -1):
        j = _rand_range(i)
        a[i], a[j] = a[j], a[i]
    return a


def _python_reference(n_dim, n_fact):
    LIBC.srand(0)
    dim_keys = _unique_shuffled(n_dim)
    fact = []
    for _ in range(n_fact // n_dim):
        fact += _unique_shuffled(n_dim)
    if n_fact % n_dim:
        fact += _unique_shuffled(n_fact % n_dim)
    return dim_keys, fact


@pytest.mark.parametrize("n_dim,n_fact", [(7, 20), (1000, 2500), (1, 3), (4096, 4096), (300, 0)])
def test_reference_generator_matches_rand_sequence(n_dim, n_fact):
    got = L.gen_reference_host(n_dim, n_fact)
    dk, fk = _python_reference(n_dim, n_fact)
    assert got["dim_keys"].tolist() == dk
    assert got["fact_keys"].tolist() == fk
    assert np.array_equal(got["dim_vals"], np.arange(n_dim, dtype=np.int32))
    assert np.array_equal(got["fact_vals"], np.arange(n_fact, dtype=np.int32))
    # every block of the fact keys is a permutation of 1..block size
    for b in range(0, n_fact, n_dim):
        blk = got["fact_keys"][b:b + n_dim]
        assert sorted(blk.tolist()) == list(range(1, blk.size + 1))


def test_reference_generator_skipped_columns_keep_the_sequence():
    a = L.gen_reference_host(500, 1700)
    b = L.gen_reference_host(500, 1700, want=("fact_keys",))
    assert np.array_equal(a["fact_keys"], b["fact_keys"])


def test_reference_generator_rejects_bad_sizes():
    with pytest.raises(L.LJError):
        L.gen_reference_host(0, 10)
