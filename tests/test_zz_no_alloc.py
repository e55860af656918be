"""lj_star_join_noalloc: the star join over resident dimensions that
allocates nothing (its buffers are a fixed per-GPU device array).

Its results must equal lj_star_join's on the same inputs -- groups, sums,
counts, the int32 sums and the exact per-stage statistics -- on two calls in
a row, the second over fewer fact rows, and the shared array must be clean
after a call that fails.
"""
import numpy as np
import pytest
import torch

import lipjoin as L
from lipjoin import workloads as W
from q33_helpers import gpu_groups, oracle_q33, to_device

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = L.Context(0)
    seed = W.seed_of(3) + 17
    n = (1 << 20) + 77
    host_dims = W.q33_host_dims(seed)
    host_fact = W.q33_host_fact(seed, n)
    dev = to_device({**host_dims, **host_fact})
    city = ((L.COMB_SET, L.CMP_LT, 20),)      # up to 2,800 groups: all listed
    dims = [ctx.build_dim(d) for d in W.q33_dim_descs(dev, city_pred=city)]
    yield ctx, n, host_dims, host_fact, dev, dims, city
    for d in dims:
        d.release()
    ctx.close()


def _same(a, b, static=True, payloads=True):
    """Equal results; the per-stage statistics too for a static plan (an
    adaptive plan's order comes from sample totals that each wave reads as
    they accumulate, so its statistics vary from run to run).  payloads
    False: a payload outside the group key is any kept row's value."""
    assert a["status"] == b["status"] == L._abi.OK
    if payloads:
        assert a["groups"] == b["groups"]
    else:
        assert [(g[0], g[2], g[3], g[4]) for g in a["groups"]] == [(g[0], g[2], g[3], g[4]) for g in b["groups"]]
    assert (a["n_groups"], a["n_rows"], a["sum_total"]) == (b["n_groups"], b["n_rows"], b["sum_total"])
    if static:
        assert a["stats"]["rows_in"] == b["stats"]["rows_in"] and a["stats"]["lines_in"] == b["stats"]["lines_in"]


@pytest.mark.parametrize("plan", ["0-lip", "3-naive", "lip-adaptive"])
def test_noalloc_equals_star_join_twice(setup, plan):
    ctx, n, host_dims, host_fact, dev, dims, city = setup
    name, method, order, bo = [p for p in W.q33_plans() if p[0] == plan][0]
    for rows in (n, (1 << 17) + 3):          # the second call over fewer rows
        _flush(f"noalloc {plan} rows={rows}")
        q = W.q33_star_desc(dev, rows, method, order, bo, resident=dims)
        a = ctx.star_join_noalloc(q)
        b = ctx.star_join(q)
        _same(a, b, static=method != "lip-adaptive")
        if rows != n:
            hf = {k: v[:rows] for k, v in host_fact.items()}
            ref, _, _, _ = oracle_q33(host_dims, hf, order, method, bo, city_pred=city)
            assert gpu_groups(a) == ref


def test_noalloc_single_group(setup):
    ctx, n, host_dims, host_fact, dev, dims, city = setup
    _flush("noalloc single group")
    for rows in ((1 << 20) + 1, 4099):
        q = L.star_desc([dev["lo_custkey"]], dev["lo_revenue"], rows, None, (0,), method="lip", resident=[dims[1]])
        _same(ctx.star_join_noalloc(q), ctx.star_join(q), payloads=False)


def test_noalloc_rejects_and_recovers(setup):
    ctx, n, host_dims, host_fact, dev, dims, city = setup
    _flush("noalloc errors")
    q = W.q33_star_desc(dev, n, "lip", W.Q33_ORDERS[0])          # dimensions built inside: refused
    with pytest.raises(L.LJError) as e:
        ctx.star_join_noalloc(q)
    assert e.value.status == L._abi.EINVAL
    big = W.q33_star_desc(dev, n, "lip", W.Q33_ORDERS[0], resident=dims)
    big.n_groups = (1 << 19) + 1
    with pytest.raises(L.LJError) as e:
        ctx.star_join_noalloc(big)
    assert e.value.status == L._abi.EINVAL
    # more groups than the array lists: refused, and the array is clean after
    wide = ((L.COMB_SET, L.CMP_LT, 120),)
    wdims = [ctx.build_dim(d) for d in W.q33_dim_descs(dev, city_pred=wide)]
    try:
        qw = W.q33_star_desc(dev, n, "lip", W.Q33_ORDERS[0], resident=wdims)
        with pytest.raises(L.LJError) as e:
            ctx.star_join_noalloc(qw)
        assert e.value.status == L._abi.EUNSUPPORTED
    finally:
        for d in wdims:
            d.release()
    q = W.q33_star_desc(dev, n, "lip", W.Q33_ORDERS[0], resident=dims)
    _same(ctx.star_join_noalloc(q), ctx.star_join(q))
