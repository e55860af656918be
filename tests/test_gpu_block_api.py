"""The block operator API (include/lipjoin/block/crystal.hpp) is a drop-in for
the reference's crystal/*.cuh: kernels written against it (tests/block_api/
block_api.hip, built by __graft_entry__.build()) run at the reference's
shapes <128,4> and <256,2> on ragged tiles, and every output is checked
against the oracle:

  build: BlockLoad, InitFlags, BlockPredGTE / AndLT / OrEQ (pred.cuh:35-356),
         BlockBuildBloomFilter (bloom.cuh:10-88), BlockBuildSelectivePHT_1/_2
         (join.cuh:191-335)            -> Bloom words and both tables, word for word
  probe: BlockLoad, BlockProbeBloomFilter (bloom.cuh:90-174), IsTerm
         (term.cuh:25-33, per thread), BlockProbeAndPHT_1/_2 (join.cuh:29-189),
         BlockPredLoad (load.cuh:25-70), BlockPredAndLTE / AndGT, BlockStore
         (store.cuh:25-69), BlockSum (reduce.cuh:25-75)
                                       -> per-row flags and payloads, per-thread IsTerm, the sum
"""
import ctypes as C
import os

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "block_api", "_build", "libblockapi.so")
SHAPES = {0: (128, 4), 1: (256, 2)}


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(LIB), "tests/block_api was not built (run __graft_entry__.build())"
    lb = C.CDLL(LIB)
    lb.ba_build.restype = C.c_int
    lb.ba_build.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                            C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    lb.ba_probe.restype = C.c_int
    lb.ba_probe.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                            C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_void_p]
    return lb


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _slots(keys, kmin, length):
    """HASH (join.cuh:27) per key, -1 below keys_min (the reference's negative remainder)."""
    x = keys.astype(np.int64) - kmin
    return np.where(x >= 0, np.mod(x, length), -1)


@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("n_dim,n_fact,kmin", [(5000, 20011, 0), (3001, 4099, 1000), (700, 513, 7)])
def test_block_api_build_and_probe(lib, shape, n_dim, n_fact, kmin):
    bt, ipt = SHAPES[shape]
    rng = np.random.default_rng(n_dim * 7 + shape)
    # dimension: unique keys kmin..kmin+n_dim-1 in shuffled order, plus keys below keys_min
    keys = np.concatenate([rng.permutation(n_dim).astype(np.int32) + kmin,
                           np.array([kmin - 1, kmin - 5], dtype=np.int32)])
    vals = rng.integers(0, 100, keys.size).astype(np.int32)
    lo, hi, eq = 10, 30, 95                                   # about 21 % of the rows selected
    ht_len, bf_words = n_dim, max(1, n_dim // 64)            # Bloom bits < key range: MOD folds keys
    steps = [(O.COMB_SET, O.CMP_GTE, lo), (O.COMB_AND, O.CMP_LT, hi), (O.COMB_OR, O.CMP_EQ, eq)]
    flags = O.predicate(vals, steps)
    spec = O.bloom_spec(O.BLOOM_MOD, bf_words * 32, 1, keys_min=kmin)
    ref_bf = O.bloom_build(spec, keys, flags)
    ref_ht1, _ = O.pht1_build(keys, flags, ht_len, kmin)
    ref_ht2, _ = O.pht2_build(keys, vals, flags, ht_len, kmin)

    dk, dv = _dev(keys), _dev(vals)
    bf = torch.zeros(bf_words, dtype=torch.int32, device="cuda")
    ht1 = torch.zeros(ht_len, dtype=torch.int32, device="cuda")
    ht2 = torch.zeros(2 * ht_len, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    assert lib.ba_build(shape, dk.data_ptr(), dv.data_ptr(), keys.size, lo, hi, eq, kmin, bf.data_ptr(), bf_words,
                        ht1.data_ptr(), ht2.data_ptr(), ht_len, stream) == 0
    torch.cuda.synchronize()
    assert np.array_equal(bf.cpu().numpy().view(np.uint32), ref_bf)
    assert np.array_equal(ht1.cpu().numpy(), ref_ht1)
    assert np.array_equal(ht2.cpu().numpy(), ref_ht2)

    # fact: keys in and around the domain (below: miss; above: aliases, no key compare)
    fk = rng.integers(kmin - 50, kmin + n_dim + 50, n_fact).astype(np.int32)
    meas = rng.integers(1, 10001, n_fact).astype(np.int32)
    mhi, mlo = 9000, 500
    tile = bt * ipt
    n_tiles = (n_fact + tile - 1) // tile
    dfk, dm = _dev(fk), _dev(meas)
    f_out = torch.full((n_fact,), -1, dtype=torch.int32, device="cuda")
    p_out = torch.full((n_fact,), -1, dtype=torch.int32, device="cuda")
    t_out = torch.full((n_tiles * bt,), -1, dtype=torch.int32, device="cuda")
    s_out = torch.zeros(1, dtype=torch.int64, device="cuda")
    assert lib.ba_probe(shape, dfk.data_ptr(), dm.data_ptr(), n_fact, bf.data_ptr(), bf_words, kmin,
                        ht1.data_ptr(), ht2.data_ptr(), ht_len, mhi, mlo, f_out.data_ptr(), p_out.data_ptr(),
                        t_out.data_ptr(), s_out.data_ptr(), stream) == 0
    torch.cuda.synchronize()

    # oracle, row by row
    bloom_pass = np.array([O.bloom_probe(spec, ref_bf, int(k)) for k in fk])
    sl = _slots(fk, kmin, ht_len)
    ok = sl >= 0
    h1 = ok & (ref_ht1[np.where(ok, sl, 0)] != 0)
    pair = ref_ht2.view(np.uint64)[np.where(ok, sl, 0)]
    h2 = ok & (pair != 0)
    pay = (pair >> np.uint64(32)).astype(np.uint32).view(np.int32)
    got_pay = bloom_pass & h1 & h2
    exp_pay = np.where(got_pay, pay, 0)
    exp_flag = (got_pay & (meas <= mhi) & (meas > mlo)).astype(np.int32)
    assert np.array_equal(f_out.cpu().numpy(), exp_flag)
    assert np.array_equal(p_out.cpu().numpy(), exp_pay)
    # IsTerm per thread after the Bloom probe: every item of the thread dead,
    # where an item past the last row keeps InitFlags' 1
    padded = np.ones(n_tiles * tile, dtype=bool)
    padded[:n_fact] = bloom_pass
    exp_term = ~padded.reshape(n_tiles, ipt, bt).any(axis=1)          # [tile, thread]
    assert np.array_equal(t_out.cpu().numpy().reshape(n_tiles, bt), exp_term.astype(np.int32))
    exp_sum = int(np.sum(np.where(exp_flag == 1, meas.astype(np.int64) * exp_pay, 0)))
    assert int(s_out.item()) == exp_sum
    assert exp_flag.sum() > 0 and exp_term.any() and not exp_term.all()
