"""Full-size parity: the configurations at the sizes they are benchmarked at,
GPU (through the C-ABI) against the CPU oracle on the same seeded inputs
(the oracle is multithreaded C; each case finishes in seconds)."""
import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O
from lipjoin import workloads as W
from q33_helpers import gpu_groups, oracle_q33, to_device

pytestmark = pytest.mark.gpu

THREADS = 16


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def q33_full(ctx):
    n = W.Q33_FACT_ROWS
    seed = W.seed_of(3)
    host_dims = W.q33_host_dims(seed)
    host_fact = W.q33_host_fact(seed, n, ["lo_suppkey", "lo_custkey", "lo_orderdate", "lo_revenue"])
    dev = W.q33_device_columns(ctx, seed, n)
    torch.cuda.synchronize()
    return n, host_dims, host_fact, dev


def test_q33_120M_device_columns_match_host(q33_full):
    n, host_dims, host_fact, dev = q33_full
    for k in ("lo_suppkey", "lo_custkey", "lo_orderdate", "lo_revenue"):
        idx = np.random.default_rng(1).integers(0, n, 4096)
        assert np.array_equal(dev[k][torch.from_numpy(idx).cuda()].cpu().numpy(), host_fact[k][idx]), k


@pytest.mark.parametrize("plan", ["0-lip", "lip-adaptive", "2-lip", "3-naive"])
def test_q33_120M_plan_vs_oracle(ctx, q33_full, plan):
    _flush(f"q33 120M {plan}")
    n, host_dims, host_fact, dev = q33_full
    name, method, order, bo = [p for p in W.q33_plans() if p[0] == plan][0]
    res = ctx.star_join(W.q33_star_desc(dev, n, method, order, bo))
    groups, lines, rows, oob = oracle_q33(host_dims, host_fact, order, method, bo, threads=THREADS)
    assert oob == 0
    assert gpu_groups(res) == groups and len(groups) == 24
    if method != "lip-adaptive":
        assert res["stats"]["lines_in"][:4] == lines
        assert res["stats"]["rows_in"][:4] == rows[:4]


def test_config1_plain_join_full(ctx):
    """2^20-row dimension (attr < 10 of U[0,99]), 2^24 fact rows, no Bloom filter;
    join.cu's SUM(measure * payload) and SUM(measure), key-0 quirk included."""
    _flush("config 1 full")
    seed, nd, nf = W.seed_of(1), 1 << 20, 1 << 24
    keys, attr, fk, meas = W.single_join_host(seed, nd, nf, 99)
    dk, da, dfk, dm = W.single_join_device(ctx, seed, nd, nf, 99)
    pred = W.single_join_pred(10)
    desc = L.dim_desc(dk, nd, nd, attr=da, predicate=pred)
    r = ctx.hash_join(dfk, dm, nf, desc, method="naive")
    ref = O.single_join(O.DimSpec(keys, attr, nd, pred=list(pred)), fk, meas, method=O.NAIVE, threads=THREADS)
    assert (r["n_match"], r["sum_measure"], r["sum_product"]) == (ref["n_match"], ref["sum_measure"], ref["sum_product"])


@pytest.mark.parametrize("bloom", [("mult", 1 << 22, 2, L.BLOOM_STANDARD), ("mod", 1 << 22)])
def test_config2_bloom_join_full(ctx, bloom):
    """2^22-row dimension, 2^28 fact rows, 10 % selectivity, LIP with a Bloom filter."""
    _flush(f"config 2 full {bloom}")
    seed, nd, nf = W.seed_of(2), 1 << 22, 1 << 28
    keys, attr, fk, meas = W.single_join_host(seed, nd, nf, 9999)
    dk, da, dfk, dm = W.single_join_device(ctx, seed, nd, nf, 9999)
    pred = W.single_join_pred(1000)
    desc = L.dim_desc(dk, nd, nd, attr=da, predicate=pred, bloom=bloom)
    r = ctx.hash_join(dfk, dm, nf, desc, method="lip")
    ob = O.bloom_spec(O.BLOOM_MOD, bloom[1], 1) if bloom[0] == "mod" else \
        O.bloom_spec(O.BLOOM_MULT, bloom[1], bloom[2], layout=bloom[3])
    ref = O.single_join(O.DimSpec(keys, attr, nd, pred=list(pred), bloom=ob), fk, meas, method=O.LIP,
                        threads=THREADS)
    assert (r["n_selected"], r["n_bloom_pass"], r["n_match"], r["sum_measure"]) == \
        (ref["n_selected"], ref["n_bloom_pass"], ref["n_match"], ref["sum_measure"])
    del dfk, dm
    torch.cuda.empty_cache()
