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


# ---- the star-schema generator (src/ops/utils/star.cuh:67-120) -------------------------
def _python_reference_star(n_fact, dim_rows):
    LIBC.srand(0)
    fact_keys = _unique_shuffled(n_fact)
    dims, fks = [], []
    for r in dim_rows:
        dims.append(_unique_shuffled(r))
        fk = []
        for _ in range(n_fact // r):
            fk += _unique_shuffled(r)
        if n_fact % r:
            fk += _unique_shuffled(n_fact % r)
        fks.append(fk)
    return fact_keys, dims, fks


@pytest.mark.parametrize("n_fact,dim_rows", [(10, [2]), (1000, [7, 300, 1000]), (257, [1000, 3, 16, 256]),
                                             (0, [5])])
def test_reference_star_generator_matches_rand_sequence(n_fact, dim_rows):
    from lipjoin.star_schema import StarSchema
    s = StarSchema(n_fact, dim_rows, [0] * len(dim_rows))
    got = s.host_columns()
    fk_ref, dims_ref, fks_ref = _python_reference_star(n_fact, dim_rows)
    assert got["fact_keys"].tolist() == fk_ref
    assert np.array_equal(got["fact_vals"], np.arange(n_fact, dtype=np.int32))
    for d, r in enumerate(dim_rows):
        assert got["dim_keys"][d].tolist() == dims_ref[d]
        assert np.array_equal(got["dim_vals"][d], np.arange(r, dtype=np.int32))
        assert got["fact_fks"][d].tolist() == fks_ref[d]


def test_reference_star_generator_null_columns_keep_the_sequence():
    import ctypes as C
    from lipjoin.star_schema import StarSchema
    rows = (C.c_int64 * 2)(50, 7)
    full = StarSchema(120, [50, 7], [0, 0]).host_columns()
    fk1 = np.empty(120, np.int32)
    st = L.lib().lj_gen_reference_star_host(120, 2, rows, None, None, None, None,
                                            (C.c_void_p * 2)(None, fk1.ctypes.data))
    assert st == 0 and np.array_equal(fk1, full["fact_fks"][1])


def test_star_schema_arguments_and_seeded_columns():
    from lipjoin.star_schema import StarSchema
    with pytest.raises(ValueError):
        StarSchema(10, [], [])
    with pytest.raises(ValueError):
        StarSchema(10, [4], [1], htable_factor=3)
    with pytest.raises(ValueError):
        StarSchema(10, [4, 4, 4, 4, 4], [1] * 5)
    a = StarSchema(5000, [100, 37], [9, 3], gen="seeded", seed=11).host_columns()
    b = StarSchema(5000, [100, 37], [9, 3], gen="seeded", seed=11).host_columns()
    assert all(np.array_equal(x, y) for x, y in zip(a["fact_fks"], b["fact_fks"]))
    assert sorted(a["dim_keys"][1].tolist()) == list(range(1, 38))
    assert a["fact_fks"][0].min() >= 1 and a["fact_fks"][0].max() <= 100
    assert sorted(a["fact_keys"].tolist()) == list(range(1, 5001))
    import ctypes as C
    assert L.lib().lj_gen_reference_star_host(10, 5, (C.c_int64 * 5)(1, 1, 1, 1, 1), None, None, None, None,
                                              None) == L._abi.EINVAL
