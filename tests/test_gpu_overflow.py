"""Group-list overflow must leave clean state behind, and one failing query
must fail alone.

A query that hits more groups than can be listed returns LJ_EUNSUPPORTED.
The finalize zeroes only the listed groups, so the library clears the
query's whole accumulator region before the next query or run.  Without that
clear, the unlisted groups keep their sums: the next query adds onto them and
never lists them (status OK, groups missing or wrong).  These tests run the
overflowing query first and then check the next results against the oracle.
"""
import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _dim(n_dim, keys, attr, pred=()):
    return L.dim_desc(keys, n_dim, n_dim, attr=attr, predicate=pred)


def _oracle_groups(keys, attr, pred, fk, meas, n_groups):
    d = O.DimSpec(keys, attr, keys.size, pred=list(pred))
    groups, _, _, _ = O.star_join([d], [fk], meas, (0,), method=O.NAIVE, group_mult=(1, 0, 0, 0),
                                  n_groups=n_groups, threads=8)
    return [(g, s, c) for (g, _, s, c) in groups]


def _gpu_groups(res):
    return [(g, s, c) for (g, _, s, c, _) in res["groups"]]


def test_context_clean_after_touched_overflow():
    """2^21 groups over 2^23 rows hit about 2.06 M groups, more than the
    2^20 the context lists; the next query on the same context must equal
    the oracle."""
    n_dim, n_fact = 1 << 21, 1 << 23
    rng = np.random.default_rng(11)
    keys = np.arange(n_dim, dtype=np.int32)
    attr = keys.copy()                         # payload = key: group g = key
    fk = rng.integers(0, n_dim, n_fact, dtype=np.int32)
    meas = rng.integers(1, 10001, n_fact, dtype=np.int32)
    dk, da, dfk, dm = (torch.from_numpy(x).cuda() for x in (keys, attr, fk, meas))
    with L.Context(0) as ctx:
        _flush("overflow: wide query")
        q = L.star_desc([dfk], dm, n_fact, [_dim(n_dim, dk, da)], (0,), method="naive",
                        group_mult=(1, 0, 0, 0), n_groups=n_dim)
        with pytest.raises(L.LJError) as e:
            ctx.star_join(q)
        assert e.value.status == L._abi.EUNSUPPORTED
        _flush("overflow: narrow query on the same context")
        pred = ((L.COMB_SET, L.CMP_LT, 3000),)
        q2 = L.star_desc([dfk], dm, n_fact, [_dim(n_dim, dk, da, pred)], (0,), method="naive",
                         group_mult=(1, 0, 0, 0), n_groups=n_dim)
        res = ctx.star_join(q2, groups_cap=4096)
        ref = _oracle_groups(keys, attr, pred, fk, meas, n_dim)
        assert res["status"] == L._abi.OK and res["n_groups"] == len(ref)
        assert _gpu_groups(res) == ref
        # and the wide query still fails the same way (not OK with stale groups)
        with pytest.raises(L.LJError):
            ctx.star_join(q)
        assert _gpu_groups(ctx.star_join(q2, groups_cap=4096)) == ref


def test_batch_overflow_fails_alone_and_again(ctx):
    """A batch query over groups_cap fails with its own status on every run;
    the other query of the batch is exact on every run."""
    n_dim, n_fact = 1 << 12, (1 << 18) + 3
    rng = np.random.default_rng(12)
    keys = np.arange(n_dim, dtype=np.int32)
    attr = rng.integers(0, 100, n_dim, dtype=np.int32)
    fk = rng.integers(0, n_dim, n_fact, dtype=np.int32)
    meas = rng.integers(1, 10001, n_fact, dtype=np.int32)
    dk, da, dfk, dm = (torch.from_numpy(x).cuda() for x in (keys, attr, fk, meas))
    wide = L.star_desc([dfk], dm, n_fact, [_dim(n_dim, dk, da)], (0,), method="naive",
                       group_mult=(1, 0, 0, 0), n_groups=100)      # group = attr: ~100 groups
    pred = ((L.COMB_SET, L.CMP_LT, 2),)
    narrow = L.star_desc([dfk], dm, n_fact, [_dim(n_dim, dk, da, pred)], (0,), method="lip",
                         group_mult=(1, 0, 0, 0), n_groups=100)     # 2 groups
    ref = _oracle_groups(keys, attr, pred, fk, meas, 100)
    ref_wide = _oracle_groups(keys, attr, (), fk, meas, 100)
    assert len(ref_wide) > 4 and len(ref) == 2
    for graph in (False, True):
        _flush(f"batch overflow, graph={graph}")
        b = ctx.batch([wide, narrow], use_graph=graph, groups_cap=4)
        for it in range(3):
            res = b.run(allow_errors=True)
            assert res[0]["status"] == L._abi.EUNSUPPORTED, (graph, it)
            assert res[0]["n_groups"] == 4
            assert res[1]["status"] == L._abi.OK and _gpu_groups(res[1]) == ref, (graph, it)
        b.close()
    # the same wide query with room for its groups is exact
    b = ctx.batch([wide], groups_cap=128)
    assert _gpu_groups(b.run()[0]) == ref_wide
    b.close()


def test_serve_bad_request_fails_alone(ctx):
    """One over-cap request in a burst fails with its own status; a malformed
    one is refused at submit; every other request is exact."""
    n_dim, n_pool = 1 << 12, 1 << 20
    rng = np.random.default_rng(13)
    keys = np.arange(n_dim, dtype=np.int32)
    attr = rng.integers(0, 100, n_dim, dtype=np.int32)
    fk = rng.integers(0, n_dim, n_pool, dtype=np.int32)
    meas = rng.integers(1, 10001, n_pool, dtype=np.int32)
    dk, da, dfk, dm = (torch.from_numpy(x).cuda() for x in (keys, attr, fk, meas))
    pred = ((L.COMB_SET, L.CMP_LT, 20),)
    dim = ctx.build_dim(L.dim_desc(dk, n_dim, n_dim, attr=da, predicate=pred, bloom=("mod", n_dim)))
    dim_all = ctx.build_dim(L.dim_desc(dk, n_dim, n_dim, attr=da))
    torch.cuda.synchronize()
    slices = [(int(o) * 4, int(n)) for o, n in zip(rng.integers(0, (n_pool - 70000) // 4, 48),
                                                   rng.integers(1, 70000, 48))]

    def desc(o, n, d=dim, ng=1, gm=(0, 0, 0, 0)):
        return L.star_desc([dfk.data_ptr() + 4 * o], dm.data_ptr() + 4 * o, n, None, (0,), method="lip",
                           resident=[d], group_mult=gm, n_groups=ng)

    srv = ctx.serve(max_batch=16, max_inflight=2, groups_cap=8)
    ids = []
    for i, (o, n) in enumerate(slices):
        ids.append(srv.submit(desc(o, n)))
        if i == 10:
            bad_cap = srv.submit(desc(o, n, dim_all, ng=100, gm=(1, 0, 0, 0)))   # ~100 groups > 8
    bad = desc(0, 100)
    bad.order[0] = 2
    with pytest.raises(L.LJError) as e:
        srv.submit(bad)
    assert e.value.status == L._abi.EINVAL and "order" in str(e.value)
    out = [srv.poll(i, timeout_ms=60000) for i in ids]
    r_bad = srv.poll(bad_cap, timeout_ms=60000)
    srv.stop()
    assert r_bad["status"] == L._abi.EUNSUPPORTED
    spec = O.DimSpec(keys, attr, n_dim, pred=list(pred))
    for (o, n), r in zip(slices, out):
        ref = O.single_join(spec, fk[o:o + n], meas[o:o + n], method=O.NAIVE, threads=8)
        assert r["status"] == L._abi.OK
        assert (r["n_rows"], r["sum_total"]) == (ref["n_match"], ref["sum_measure"])
    dim.release()
    dim_all.release()
