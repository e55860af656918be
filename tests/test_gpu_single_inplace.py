"""One-dimension, one-group LIP joins summed in place.

When the Bloom filter passes enough rows for the measure column to stream
with the keys, a single join over a perfect-hash table (PHT_1 or PHT_2)
probes the table and adds each survivor's measure and product in place
instead of queueing it (LJ_DIRECT, on by default).  Every result must equal
the queued path's (a context created with LJ_DIRECT=0) and the oracle's:
match count, SUM(measure), SUM(measure * payload) (where no two selected
keys share a table slot), filter pass count and,
for the static plan, the per-stage row and line counts.  The cases cover
both table kinds, a table shorter than the key range (slots alias, as the
reference's modulo hash does), pass rates on either side of the streaming
threshold, a lowered threshold (LJ_DIRECT_EAGER, with LJ_DIRECT=2: the
in-place body even where the filter's selected keys predict few rows
passing), the table's occupancy
bitmap in LDS deciding the hits (LJ_DIRECT_OCC=2; a slot count that is not a
multiple of 64 in the aliased case) with the measure streamed or, below
1/8 of the rows hitting (LJ_DIRECT_OCC_EAGER), gathered for the hits only,
the adaptive plan, and ragged and tiny fact tables.
"""
import os

import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O
from lipjoin import workloads as W
from q33_helpers import to_device

pytestmark = pytest.mark.gpu

SEED = W.seed_of(7)


def _flush(msg):
    print(msg, flush=True)


def _context_with(env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return L.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def ctx_queue():
    c = _context_with({"LJ_DIRECT": "0"})
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_low_threshold():
    # the in-place body whatever the filter's estimated pass rate, streaming
    # from 1/32 passing
    c = _context_with({"LJ_DIRECT": "2", "LJ_DIRECT_EAGER": "32"})
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_occ():
    # the table's occupancy bitmap in LDS whenever it fits (the default needs
    # about CUs x table slots fact rows)
    c = _context_with({"LJ_DIRECT_OCC": "2"})
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_occ_low():
    # ... and the hits' measures gathered, never streamed (streamed above
    # 1/n of the rows hitting: n = 1 never streams)
    c = _context_with({"LJ_DIRECT": "2", "LJ_DIRECT_OCC": "2", "LJ_DIRECT_EAGER": "32",
                       "LJ_DIRECT_OCC_EAGER": "1"})
    yield c
    c.close()


# (n_dim, n_fact, ht_len divisor, threshold on attr (of 0..99), filter bits, table kind)
CASES = [
    (1 << 16, (1 << 20) + 13, 1, 50, 1 << 14, L.HT_PHT2),   # passes ~85 %: streamed, in place
    (1 << 16, (1 << 20) + 13, 1, 50, 1 << 14, L.HT_PHT1),   # keys-only table: products are 0
    (1 << 16, (1 << 20) + 7, 3, 40, 1 << 15, L.HT_PHT2),    # table of n_dim/3 slots: keys alias
    (1 << 16, (1 << 20) + 1, 1, 5, 1 << 20, L.HT_PHT2),     # passes ~5 %: the queue
    (1 << 16, (1 << 20) + 3, 1, 12, 1 << 17, L.HT_PHT2),    # passes ~12 %: queue, or in place when lowered
    (1 << 16, (1 << 20) + 9, 1, 2, 1 << 12, L.HT_PHT2),     # passes ~27 %, 2 % hit: measures gathered
    (1 << 12, (1 << 20) + 11, 1, 60, 1 << 11, L.HT_PHT2),   # small table: the default builds its bitmap
    (1 << 12, 4099, 1, 70, 1 << 10, L.HT_PHT2),              # tiny, ragged
    (1 << 12, 100, 1, 70, 1 << 10, L.HT_PHT2),               # fewer rows than one tile
]


def _run(c, dev, n_dim, n_fact, ht_len, thr, bits, kind, method):
    desc = L.dim_desc(dev["keys"], n_dim, ht_len, attr=dev["attr"], predicate=W.single_join_pred(thr),
                      bloom=("mod", bits), ht_kind=kind)
    return c.hash_join(dev["fk"], dev["meas"], n_fact, desc, method=method)


def _key(r, product=True):
    return (r["n_match"], r["sum_measure"], r["sum_product"] if product else None, r["n_bloom_pass"])


@pytest.mark.parametrize("method", ["lip", "lip-adaptive"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_in_place_equals_queue_and_oracle(ctx, ctx_queue, ctx_low_threshold, ctx_occ, ctx_occ_low, case, method):
    n_dim, n_fact, div, thr, bits, kind = CASES[case]
    ht_len = n_dim // div
    _flush(f"in-place single join case {case} {method}: n_dim {n_dim} n_fact {n_fact} ht_len {ht_len} "
           f"thr {thr} bits {bits} kind {kind}")
    keys, attr, fk, meas = W.single_join_host(SEED + case, n_dim, n_fact, 99)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    ref = O.single_join(O.DimSpec(keys, attr, ht_len, pred=list(W.single_join_pred(thr)), ht_kind=kind,
                                  bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1)), fk, meas, method=O.LIP, threads=8)
    # two selected keys in one slot (an aliased table): the slot keeps the
    # payload of whichever row the build wrote first, so the product is not
    # deterministic -- the reference's CAS build has the same race
    product = div == 1
    exp = (ref["n_match"], ref["sum_measure"], ref["sum_product"] if product else None, ref["n_bloom_pass"])
    results = [_run(c, dev, n_dim, n_fact, ht_len, thr, bits, kind, method)
               for c in (ctx, ctx_queue, ctx_low_threshold, ctx_occ, ctx_occ_low)]
    torch.cuda.synchronize()
    for i, r in enumerate(results):
        assert _key(r, product) == exp, (i, _key(r, product), exp)
        if kind == L.HT_PHT1:
            assert r["sum_product"] == 0
        if method == "lip":
            st = r["stats"]
            assert st["rows_in"][:2] == [n_fact, ref["n_match"]], i
            assert st["lines_in"][:2] == [ref["lines_fk"], ref["lines_measure"]], i


@pytest.mark.parametrize("n_dim,thr", [(1 << 16, 5), (1 << 16, 50), (1 << 12, 50)])
def test_resident_dimension_matches_inline(ctx, n_dim, thr):
    """The in-place body and the table's occupancy bitmap are chosen from the
    filter's selected keys: on the host for a resident dimension, on the
    device (the build's counters) for one built with the query.  At 5 %
    selected (a 2^14-bit filter over 2^16 keys: about 18 % of the fact rows
    pass) neither chooses the body, at 50 % both do; over 2^12 keys both
    also build the bitmap.  Either way the one-group star join equals the
    oracle's."""
    n_fact = (1 << 20) + 5
    _flush(f"resident vs inline one-dimension star join, {n_dim} keys, threshold {thr}")
    keys, attr, fk, meas = W.single_join_host(SEED + 20 + thr, n_dim, n_fact, 99)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    pred = W.single_join_pred(thr)
    desc = L.dim_desc(dev["keys"], n_dim, n_dim, attr=dev["attr"], predicate=pred, bloom=("mod", 1 << 14))
    dim = ctx.build_dim(desc)
    try:
        a = ctx.star_join(L.star_desc([dev["fk"]], dev["meas"], n_fact, None, [0], resident=[dim]))
        b = ctx.star_join(L.star_desc([dev["fk"]], dev["meas"], n_fact, [desc], [0]))
    finally:
        dim.release()
    ref = O.single_join(O.DimSpec(keys, attr, n_dim, pred=list(pred), bloom=O.bloom_spec(O.BLOOM_MOD, 1 << 14, 1)),
                        fk, meas, method=O.LIP, threads=8)
    exp = (ref["n_match"], ref["sum_measure"])
    assert (a["n_rows"], a["sum_total"]) == exp and (b["n_rows"], b["sum_total"]) == exp
    for r in (a, b):
        assert r["stats"]["rows_in"][:2] == [n_fact, ref["n_match"]]
        assert r["stats"]["lines_in"][:2] == [ref["lines_fk"], ref["lines_measure"]]
        assert r["stats"]["rows_bloom_pass"] == ref["n_bloom_pass"]


@pytest.mark.parametrize("per_kind", [False, True], ids=["one-launch", "launch-per-kind"])
def test_batch_in_place_matches_alone(per_kind):
    """One-dimension joins probed in place with the occupancy bitmap, as
    instances of one batch (the batched probe kernels, one launch for
    two stage-0 kinds or one per kind, eager and graph-replayed): each
    result equals the same query run alone and the oracle's."""
    _flush(f"batch of in-place single joins (per_kind={per_kind})")
    c = _context_with({"LJ_DIRECT": "2", "LJ_DIRECT_OCC": "2"})
    try:
        # (n_dim, n_fact, threshold, filter bits): power-of-two and other bit counts
        shapes = [(1 << 12, 300_007, 60, 1 << 11), (1 << 14, 500_000, 30, 3000),
                  (1 << 13, 1 << 18, 5, 1 << 12), (1 << 12, 70_001, 90, 1500)]
        descs, refs, keep = [], [], []
        for i, (n_dim, n_fact, thr, bits) in enumerate(shapes):
            keys, attr, fk, meas = W.single_join_host(SEED + 40 + i, n_dim, n_fact, 99)
            dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
            pred = W.single_join_pred(thr)
            d = L.dim_desc(dev["keys"], n_dim, n_dim, attr=dev["attr"], predicate=pred, bloom=("mod", bits))
            descs.append(L.star_desc([dev["fk"]], dev["meas"], n_fact, [d], [0]))
            keep.append(dev)
            ref = O.single_join(O.DimSpec(keys, attr, n_dim, pred=list(pred), bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1)),
                                fk, meas, method=O.LIP, threads=8)
            refs.append((ref["n_match"], ref["sum_measure"]))
        alone = [c.star_join(q) for q in descs]
        assert [(r["n_rows"], r["sum_total"]) for r in alone] == refs
        for graph in (False, True):
            b = c.batch(descs, use_graph=graph, launch_per_kind=per_kind)
            for _ in range(2):
                res = b.run()
                assert [(r["n_rows"], r["sum_total"]) for r in res] == refs, graph
                assert [r["stats"]["rows_in"][:2] for r in res] == [r["stats"]["rows_in"][:2] for r in alone]
            b.close()
        torch.cuda.synchronize()
    finally:
        c.close()


def test_serve_in_place_bitmap():
    """The serving loop (resident dimensions, continuous batching) over
    one-dimension joins probed in place with the occupancy bitmap: every
    request's count and sum equal the oracle's over its fact slice."""
    _flush("serve in-place single joins")
    c = _context_with({"LJ_DIRECT": "2", "LJ_DIRECT_OCC": "2"})
    dims = []
    try:
        n_dim, n_pool = 1 << 14, 1 << 20
        keys, attr, fk, meas = W.single_join_host(SEED + 60, n_dim, n_pool, 99)
        dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
        specs = [(40, 1 << 12), (15, 3000)]
        for thr, bits in specs:
            d = L.dim_desc(dev["keys"], n_dim, n_dim, attr=dev["attr"], predicate=W.single_join_pred(thr),
                           bloom=("mod", bits))
            dims.append(c.build_dim(d))
        rng = np.random.default_rng(SEED + 61)
        reqs = []
        for k in range(24):
            rows = int(rng.integers(1000, 1 << 18))
            off = int(rng.integers(0, (n_pool - rows) // 4 + 1)) * 4
            reqs.append((k % len(specs), off, rows))
        descs = [L.star_desc([dev["fk"][off:off + rows]], dev["meas"][off:off + rows], rows, None, [0],
                             resident=[dims[t]]) for t, off, rows in reqs]
        srv = c.serve(max_batch=8, max_inflight=2)
        ids = [srv.submit(q) for q in descs]
        out = [srv.poll(i, timeout_ms=60000) for i in ids]
        srv.stop()
        for (t, off, rows), o in zip(reqs, out):
            thr, bits = specs[t]
            ref = O.single_join(O.DimSpec(keys, attr, n_dim, pred=list(W.single_join_pred(thr)),
                                          bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1)),
                                fk[off:off + rows], meas[off:off + rows], method=O.LIP, threads=8)
            assert o is not None and o["status"] == L._abi.OK
            assert (o["n_rows"], o["sum_total"]) == (ref["n_match"], ref["sum_measure"]), (t, off, rows)
        torch.cuda.synchronize()
    finally:
        for d in dims:
            d.release()
        c.close()
