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


@pytest.fixture(scope="module")
def config2(ctx):
    """The Bloom-join configuration: 2^22-row dimension (dense keys in shuffled
    row order, attr ~ U[0,9999]), 2^28 fact rows; host and device columns."""
    seed, nd, nf = W.seed_of(2), 1 << 22, 1 << 28
    host = W.single_join_host(seed, nd, nf, 9999)
    dev = W.single_join_device(ctx, seed, nd, nf, 9999)
    torch.cuda.synchronize()
    yield nd, nf, host, dev
    del dev
    torch.cuda.empty_cache()


def _obloom(bloom):
    if bloom is None:
        return O.bloom_spec()
    if bloom[0] == "mod":
        return O.bloom_spec(O.BLOOM_MOD, bloom[1], 1)
    return O.bloom_spec(O.BLOOM_MULT, bloom[1], bloom[2], layout=bloom[3] if len(bloom) > 3 else 0)


def test_config2_filter_words_every_size_and_k(ctx, config2):
    """Bit-exact filter words of the 2^22-key dimension at 10 % selectivity for
    every filter of the configuration's sweep: m = 2^16..2^26 bits x k = 1..3
    (multiplicative, standard), the blocked layout and the reference's modulo
    hash (bloom_join.cu:128-142 sizes the filter; bloom.cuh:10-88 builds it)."""
    nd, nf, (keys, attr, _, _), (dk, da, _, _) = config2
    pred = W.single_join_pred(1000)
    flags = O.predicate(attr, list(pred))
    cases = [("mult", 1 << lg, k, L.BLOOM_STANDARD) for lg in range(16, 27) for k in (1, 2, 3)]
    cases += [("mult", 1 << lg, k, L.BLOOM_BLOCKED) for lg in (16, 20, 24, 26) for k in (2, 3)]
    cases += [("mod", 1 << 22), ("mod", 3_000_000), ("mod", 1 << 26)]
    _flush(f"config 2 filter words: {len(cases)} filters")
    for bloom in cases:
        dim = ctx.build_dim(L.dim_desc(dk, nd, nd, attr=da, predicate=pred, bloom=bloom))
        _, words = dim.export()
        info = dim.info()
        dim.release()
        ref = O.bloom_build(_obloom(bloom), keys, flags)
        assert np.array_equal(words, ref), bloom
        assert info["bloom_bits_set"] == int(np.unpackbits(ref.view(np.uint8)).sum()), bloom
        assert info["n_selected"] == int(flags.sum())


@pytest.mark.parametrize("sel", [1, 5, 10, 25, 50])
def test_config2_every_selectivity(ctx, config2, sel):
    """Survivor counts and int64 sums at 2^28 rows for every selectivity of
    the configuration, LIP with the autotuner's filter and naive."""
    nd, nf, (keys, attr, fk, meas), (dk, da, dfk, dm) = config2
    thr = 100 * sel
    pred = W.single_join_pred(thr)
    bloom, _ = W.tuned_bloom(ctx, {"n_dim": nd, "threshold": thr})
    _flush(f"config 2 full s={sel}% {bloom}")
    r = ctx.hash_join(dfk, dm, nf, L.dim_desc(dk, nd, nd, attr=da, predicate=pred, bloom=bloom), method="lip")
    ref = O.single_join(O.DimSpec(keys, attr, nd, pred=list(pred), bloom=_obloom(bloom)), fk, meas, method=O.LIP,
                        threads=THREADS)
    assert (r["n_selected"], r["n_bloom_pass"], r["n_match"], r["sum_measure"], r["sum_product"]) == \
        (ref["n_selected"], ref["n_bloom_pass"], ref["n_match"], ref["sum_measure"], ref["sum_product"])
    # SUM(measure) only, the filter tuned for a semi-join (an exact modulo
    # filter decides the table probe of in-domain keys)
    bs, _ = W.tuned_bloom(ctx, {"n_dim": nd, "threshold": thr}, semi_join=True)
    rs = ctx.hash_join(dfk, dm, nf, L.dim_desc(dk, nd, nd, attr=da, predicate=pred, bloom=bs), method="lip",
                       sum_only=True)
    refs = ref if bs == bloom else O.single_join(O.DimSpec(keys, attr, nd, pred=list(pred), bloom=_obloom(bs)), fk,
                                                 meas, method=O.LIP, threads=THREADS)
    assert (rs["n_bloom_pass"], rs["n_match"], rs["sum_measure"], rs["sum_product"]) == \
        (refs["n_bloom_pass"], refs["n_match"], refs["sum_measure"], 0)
    if sel in (1, 50):
        rn = ctx.hash_join(dfk, dm, nf, L.dim_desc(dk, nd, nd, attr=da, predicate=pred), method="naive")
        assert (rn["n_match"], rn["sum_measure"], rn["sum_product"]) == \
            (ref["n_match"], ref["sum_measure"], ref["sum_product"])
