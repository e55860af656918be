"""Batched mode (lj_batch_*) and the Bloom autotuner, through the C-ABI.

A batch runs many independent queries as one build launch and one probe
launch per stage-0 kernel kind (optionally recorded once as a HIP graph and
replayed).  Every query's result must be identical to the same query run
alone (lj_star_join) and to the CPU oracle: int64 sums, counts, group keys
and the exact per-stage statistics of the static plans.
"""
import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O
from lipjoin import workloads as W
from q33_helpers import gpu_groups, oracle_q33, to_device

pytestmark = pytest.mark.gpu

SEED = W.seed_of(4)

# (method, bloom) per instance: LDS and L2 filters, MOD / MULT standard / blocked,
# no filter, naive and adaptive -- several stage-0 kernel kinds in one batch
MIX = [("lip", ("mod", 1 << 14)), ("lip", ("mult", 1 << 16, 2, L.BLOOM_STANDARD)),
       ("lip", ("mult", 1 << 18, 3, L.BLOOM_BLOCKED)), ("lip", ("mult", 1 << 22, 1, L.BLOOM_STANDARD)),
       ("naive", None), ("lip-adaptive", ("mod", 1 << 14)), ("lip", None), ("lip", ("mod", 1 << 14))]


def _flush(msg):
    print(msg, flush=True)


def _oracle_bloom(b):
    if b is None:
        return O.bloom_spec()
    if b[0] == "mod":
        return O.bloom_spec(O.BLOOM_MOD, b[1], 1)
    return O.bloom_spec(O.BLOOM_MULT, b[1], b[2], layout=b[3])


@pytest.fixture(scope="module")
def small_batch(ctx):
    specs = W.batch_specs(SEED, n_inst=len(MIX), log_lo=10, log_span=7, n_dim=1 << 14, s_lo=0.5, s_hi=20)
    specs[2]["n_fact"] = 0          # an empty instance
    specs[3]["n_fact"] = 1000003    # ragged, larger
    cols = [W.batch_instance_device(ctx, sp) for sp in specs]
    torch.cuda.synchronize()
    descs = [W.batch_desc(c, sp, method=m, bloom=b) for c, sp, (m, b) in zip(cols, specs, MIX)]
    return specs, cols, descs


def _key(r):
    return (r["sum_total"], r["n_rows"], r["n_groups"])


def _static_stats(r):
    st = r["stats"]
    return (st["rows_in"], st["lines_in"], st["rows_bloom_pass"])


@pytest.mark.parametrize("per_kind", [False, True], ids=["one-launch", "launch-per-kind"])
def test_batch_equals_single_and_oracle(ctx, small_batch, per_kind):
    """Eight instances of six stage-0 kernel kinds: one probe launch (the
    mixed-kind kernel) or one per kind, each run three times."""
    specs, cols, descs = small_batch
    _flush(f"batch eager vs single vs oracle (per_kind={per_kind})")
    singles = [ctx.star_join(d) for d in descs]
    b = ctx.batch(descs, launch_per_kind=per_kind)
    for it in range(3):
        res = b.run()
        for i, (r, s) in enumerate(zip(res, singles)):
            assert _key(r) == _key(s), (it, i)
            if MIX[i][0] != "lip-adaptive":
                assert _static_stats(r) == _static_stats(s), (it, i)
    b.close()
    for i, (sp, (m, bl)) in enumerate(zip(specs, MIX)):
        keys, attr, fk, meas = W.batch_instance_host(sp)
        ref = O.single_join(O.DimSpec(keys, attr, sp["n_dim"], pred=list(W.single_join_pred(sp["threshold"])),
                                      bloom=_oracle_bloom(bl)),
                            fk, meas, method=O.NAIVE if m == "naive" else O.LIP, threads=8)
        assert (singles[i]["n_rows"], singles[i]["sum_total"]) == (ref["n_match"], ref["sum_measure"]), i


def test_batch_graph_replays(ctx, small_batch):
    specs, cols, descs = small_batch
    _flush("batch graph replays")
    singles = [ctx.star_join(d) for d in descs]
    b = ctx.batch(descs, use_graph=True)
    for it in range(4):
        res = b.run()
        for i, (r, s) in enumerate(zip(res, singles)):
            assert _key(r) == _key(s), (it, i)
        # the program's own work between replays
        assert _key(ctx.star_join(descs[0])) == _key(singles[0])
        torch.cuda.synchronize()
    b.close()


def test_batch_mixed_star_and_single(ctx, small_batch):
    """Q3.3-shaped star joins (groups) and single joins in one batch."""
    specs, cols, descs = small_batch
    _flush("batch q33 + single joins")
    n = (1 << 18) + 5
    host_dims = W.q33_host_dims(W.seed_of(3))
    host_fact = W.q33_host_fact(W.seed_of(3), n)
    city = ((L.COMB_SET, L.CMP_LT, 60),)
    dev = to_device({**host_dims, **host_fact})
    q_descs, refs = [], []
    for plan in ("0-lip", "2-naive", "lip-adaptive"):
        name, method, order, bo = [p for p in W.q33_plans() if p[0] == plan][0]
        dims = W.q33_dim_descs(dev, city_pred=city)
        q_descs.append(W.q33_star_desc(dev, n, method, order, bo, dims=dims))
        refs.append(oracle_q33(host_dims, host_fact, order, method, bo, city_pred=city)[0])
    all_descs = q_descs + descs[:3]
    b = ctx.batch(all_descs, groups_cap=1 << 14)
    res = b.run()
    for i in range(len(q_descs)):
        assert gpu_groups(res[i]) == refs[i], i
    for j in range(3):
        assert _key(res[len(q_descs) + j]) == _key(ctx.star_join(descs[j]))
    b.close()


def test_batch_rejects_bad_query(ctx, small_batch):
    specs, cols, descs = small_batch
    bad = W.batch_desc(cols[0], specs[0])
    bad.order[0] = 3
    with pytest.raises(L.LJError) as e:
        ctx.batch([descs[0], bad])
    assert "query 1" in str(e.value)


def test_tune_bloom_model(ctx):
    # a tiny key set: a small LDS filter with a negligible false-positive rate
    ch = ctx.tune_bloom(1000, 1 << 22, 0.001)
    assert ch["in_lds"] == 1 and ch["fpr"] < 1e-3
    # dense keys, many selected: the exact modulo filter over the key range wins
    ch = ctx.tune_bloom(800000, 1 << 22, 0.19)
    assert ch["fpr"] <= 0.05
    # the chosen cost is the minimum the model sees (monotone in fact_pass)
    a = ctx.tune_bloom(100000, 1 << 22, 0.02)["cost"]
    b = ctx.tune_bloom(100000, 1 << 22, 0.10)["cost"]
    assert a < b


def test_tune_calibrate(ctx):
    _flush("tune calibrate")
    ctx.tune_calibrate()
    ch = ctx.tune_bloom(200000, 1 << 22, 0.05)
    assert ch["bloom_bits"] >= 1 << 16 and 0.0 <= ch["fpr"] < 1.0


def test_tune_model_get_set():
    """A calibrated model carried to another context gives the same choices;
    a model outside its bounds is refused and leaves the context's as it was."""
    with L.Context(0) as a, L.Context(0) as b:
        a.tune_calibrate()
        m = a.tune_model()
        assert m["calibrated"] == 1 and m["scan_l2"] >= 1.0 and m["scan_big"] >= m["scan_l2"]
        b.set_tune_model(m)
        got = b.tune_model()
        assert all(got[k] == m[k] for k in ("scan_l2", "scan_big", "probe_extra", "line", "gather", "scan_mod",
                                            "scan_blocked", "gather_queue", "gather_lds",
                                            "gather_queue_lds"))
        for args in [(200000, 1 << 22, 0.05, False), (1 << 21, 1 << 22, 0.3, True), (5000, 1 << 20, 0.01, False)]:
            assert a.tune_bloom(*args[:3], semi_join=args[3]) == b.tune_bloom(*args[:3], semi_join=args[3])
        for bad in [dict(m, scan_l2=0.5), dict(m, scan_big=m["scan_l2"] - 0.1), dict(m, gather=-1.0),
                    dict(m, line=float("nan")), dict(m, scan_mod=0.0), dict(m, gather_queue=-2.0)]:
            with pytest.raises(L.LJError):
                b.set_tune_model(bad)
        assert b.tune_model() == got


def test_tune_bloom_loaded_filters(ctx):
    # half of 2^22 keys selected: a filter with fewer bits than keys passes
    # nearly every row, so it must not be chosen (its modelled FPR is ~1)
    ch = ctx.tune_bloom(1 << 21, 1 << 22, 0.5)
    assert ch["bloom_bits"] >= 1 << 21 and ch["fpr"] < 0.5


def test_batch_config4_full_shape_vs_oracle(ctx):
    """The batched configuration at its own shape: 64 instances of 2^20..2^27
    fact rows (log-uniform), each with its own 2^22-key dimension and
    selectivity in [0.5, 20] %, the Bloom filter from the autotuner; one batch
    (graph replay, run twice) checked instance by instance against the
    oracle's single join."""
    specs = W.batch_specs(SEED)
    assert len(specs) == 64 and specs[0]["n_dim"] == 1 << 22
    _flush(f"config 4 full shape: {sum(s['n_fact'] for s in specs)} fact rows")
    cols, descs, blooms = [], [], []
    for sp in specs:
        c = W.batch_instance_device(ctx, sp)
        bloom, _ = W.tuned_bloom(ctx, sp, semi_join=True)
        cols.append(c)
        blooms.append(bloom)
        descs.append(W.batch_desc(c, sp, method="lip", bloom=bloom))
    torch.cuda.synchronize()
    b = ctx.batch(descs, use_graph=True)
    res = b.run()
    res2 = b.run()
    b.close()
    _flush("config 4 full shape: oracle")
    for i, (sp, bl) in enumerate(zip(specs, blooms)):
        keys, attr, fk, meas = W.batch_instance_host(sp)
        ref = O.single_join(O.DimSpec(keys, attr, sp["n_dim"], pred=list(W.single_join_pred(sp["threshold"])),
                                      bloom=_oracle_bloom(bl)), fk, meas, method=O.LIP, threads=16)
        exp = (ref["n_match"], ref["sum_measure"], ref["n_bloom_pass"])
        for r in (res[i], res2[i]):
            assert r["status"] == L._abi.OK
            assert (r["n_rows"], r["sum_total"], r["stats"]["rows_bloom_pass"]) == exp, (i, sp)
    del cols
    torch.cuda.empty_cache()
