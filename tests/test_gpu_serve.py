"""Serving (lj_serve_*): continuous batching over resident dimensions.

Every request's result must be identical (int64 sum, row count, statistics
of the static plan) whether it ran alone (lj_star_join) or in the mix of a
serving run with many requests in flight; the HBM budget admission must
reject a request that can never fit and still serve the others.
"""
import threading

import numpy as np
import pytest
import torch

import lipjoin as L
from lipjoin import workloads as W

pytestmark = pytest.mark.gpu

SEED = W.seed_of(5)


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def serve_setup(ctx):
    cols, tmpls = W.serve_templates(ctx, SEED, n_dim=1 << 16, n_templates=8)
    pool = W.serve_pool(ctx, SEED, 1 << 22, n_dim=1 << 16)
    torch.cuda.synchronize()
    reqs = W.serve_requests(SEED, 300, len(tmpls), 1 << 22, log_lo=8, log_span=12)
    descs = [W.serve_desc(pool, tmpls[t], o, n) for (t, o, n) in reqs]
    alone = [ctx.star_join(d) for d in descs]
    return cols, tmpls, pool, reqs, descs, alone


def _key(r):
    return (r["sum_total"], r["n_rows"], r["stats"]["rows_in"][:2], r["stats"]["lines_in"][:2])


def test_serve_same_alone_and_in_mix(ctx, serve_setup):
    cols, tmpls, pool, reqs, descs, alone = serve_setup
    _flush("serve burst")
    srv = ctx.serve(max_batch=32, max_inflight=2)
    ids = [srv.submit(d) for d in descs]
    out = [srv.poll(i, timeout_ms=60000) for i in ids]
    st = srv.stats()
    srv.stop()
    assert all(o is not None and o["status"] == L._abi.OK for o in out)
    for i, (o, a) in enumerate(zip(out, alone)):
        assert _key(o) == _key(a), i
    assert st["completed"] == len(descs) and st["launches"] < len(descs)   # requests were batched


def test_serve_concurrent_submitters(ctx, serve_setup):
    cols, tmpls, pool, reqs, descs, alone = serve_setup
    _flush("serve threads")
    srv = ctx.serve(max_batch=16, max_inflight=3, streams=3)   # launches overlap on the device
    results = {}

    def client(k):
        for i in range(k, len(descs), 4):
            rid = srv.submit(descs[i])
            results[i] = srv.poll(rid, timeout_ms=60000)

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.stop()
    for i, a in enumerate(alone):
        assert _key(results[i]) == _key(a), i


def test_serve_budget_admission(ctx, serve_setup):
    cols, tmpls, pool, reqs, descs, alone = serve_setup
    _flush("serve budget")
    # a budget that admits one small request at a time
    big = L.star_desc([pool["fk"]], pool["meas"], 1 << 12, None, (0,), method="lip", resident=[tmpls[0]["dim"]])
    big.n_groups = 1 << 24                           # needs ~ 700 MB of accumulators
    srv = ctx.serve(hbm_budget=64 << 20, max_batch=8)
    rid_big = srv.submit(big)
    ids = [srv.submit(d) for d in descs[:40]]
    r_big = srv.poll(rid_big, timeout_ms=60000)
    out = [srv.poll(i, timeout_ms=60000) for i in ids]
    st = srv.stats()
    srv.stop()
    assert r_big["status"] == L._abi.EBUDGET
    for o, a in zip(out, alone[:40]):
        assert o["status"] == L._abi.OK and _key(o) == _key(a)
    assert st["rejected"] == 1 and st["peak_bytes_in_flight"] <= 64 << 20


def test_serve_requires_resident_dims(ctx, serve_setup):
    cols, tmpls, pool, reqs, descs, alone = serve_setup
    d = L.dim_desc(cols["keys"], 1 << 16, 1 << 16, attr=cols["attr"])
    q = L.star_desc([pool["fk"]], pool["meas"], 1024, [d], (0,))
    srv = ctx.serve()
    with pytest.raises(L.LJError):
        srv.submit(q)
    srv.stop()


def _oracle_bloom(b):
    import oracle as O
    if b is None:
        return O.bloom_spec()
    if b[0] == "mod":
        return O.bloom_spec(O.BLOOM_MOD, b[1], 1)
    return O.bloom_spec(O.BLOOM_MULT, b[1], b[2], layout=b[3])


def test_serve_config5_shape_against_oracle(ctx):
    """The serving configuration at its own shape: 8 predicate templates over
    a resident 2^22-key dimension (PHT_2 + the autotuner's Bloom filter each),
    requests of 2^18..2^26 rows sliced from a resident 2^26-row pool, Poisson
    arrivals, a 200 GB budget.  Every served result (rows, int64 sum, Bloom
    passes) must equal the oracle's single join on the same slice and
    template."""
    import oracle as O
    seed = W.seed_of(5) + 100
    n_pool = 1 << 26
    _flush("serve config-5 shape: templates and pool")
    cols, tmpls = W.serve_templates(ctx, seed)
    pool = W.serve_pool(ctx, seed, n_pool)
    torch.cuda.synchronize()
    reqs = W.serve_requests(seed, 160, len(tmpls), n_pool)
    descs = [W.serve_desc(pool, tmpls[t], o, n) for (t, o, n) in reqs]
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / 5000.0, len(descs)))
    _flush("serve config-5 shape: Poisson arrivals at 5000 req/s")
    import time
    srv = ctx.serve(hbm_budget=200 << 30, max_batch=64, max_inflight=2)
    ids = []
    t0 = time.perf_counter()
    for i, d in enumerate(descs):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        ids.append(srv.submit(d))
    out = [srv.poll(i, timeout_ms=120000) for i in ids]
    st = srv.stats()
    srv.stop()
    assert st["completed"] == len(descs) and st["rejected"] == 0
    _flush("serve config-5 shape: oracle")
    keys = cols["keys"].cpu().numpy()
    attr = cols["attr"].cpu().numpy()
    fk = pool["fk"][:n_pool].cpu().numpy()
    meas = pool["meas"][:n_pool].cpu().numpy()
    n_dim = keys.size
    specs = [O.DimSpec(keys, attr, n_dim, pred=list(W.single_join_pred(tm["threshold"])),
                       bloom=_oracle_bloom(tm["bloom"])) for tm in tmpls]
    for i, ((t, o, n), r) in enumerate(zip(reqs, out)):
        assert r is not None and r["status"] == L._abi.OK, i
        ref = O.single_join(specs[t], fk[o:o + n], meas[o:o + n], method=O.LIP, threads=16)
        assert (r["n_rows"], r["sum_total"], r["stats"]["rows_bloom_pass"]) == \
            (ref["n_match"], ref["sum_measure"], ref["n_bloom_pass"]), (i, t, o, n)
    for tm in tmpls:
        tm["dim"].release()


@pytest.mark.parametrize("warm", [1, 1000])
def test_serve_prewarmed_buffers(ctx, serve_setup, warm):
    """Request buffers allocated and cleared at start (prewarm_groups): the
    requests of that class take them, others allocate as before; every result
    is the same as alone, and a group space outside [0, 2^26] is refused."""
    cols, tmpls, pool, reqs, descs, alone = serve_setup
    _flush(f"serve prewarm {warm}")
    srv = ctx.serve(hbm_budget=1 << 30, max_batch=8, max_inflight=2, prewarm_groups=warm)
    ids = [srv.submit(d) for d in descs]
    out = [srv.poll(i, timeout_ms=60000) for i in ids]
    srv.stop()
    assert all(o is not None and o["status"] == L._abi.OK for o in out)
    for i, (o, a) in enumerate(zip(out, alone)):
        assert _key(o) == _key(a), i
    with pytest.raises(L.LJError):
        ctx.serve(prewarm_groups=-1)
    with pytest.raises(L.LJError):
        ctx.serve(prewarm_groups=(1 << 26) + 1)
