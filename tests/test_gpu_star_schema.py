"""Star joins over synthetic N-dimension star schemas (lipjoin.star_schema,
modelled on the reference's star_schema_t, src/ops/utils/star.cuh:25-64):
one to four dimensions, the reference's own generator or the seeded one,
hash-table and filter sizing factors, every method; each result (row count,
SUM of the fact values, the static plans' per-stage row and line counts)
against the oracle's star join of the same columns, and a batch of schemas
against the same queries run alone."""
import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O
from lipjoin.star_schema import StarSchema

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _oracle(schema, cols, method, order):
    dims = [O.DimSpec(cols["dim_keys"][d], cols["dim_vals"][d], schema.ht_len(d),
                      pred=[(O.COMB_SET, O.CMP_LTE, schema.dim_sels[d])], payload=cols["dim_vals"][d], keys_min=0,
                      bloom=(O.bloom_spec(O.BLOOM_MOD, schema.bloom_bits(d), 1) if method != "naive" else None))
            for d in range(schema.n_dims)]
    groups, lines, rows, _ = O.star_join(dims, cols["fact_fks"], cols["fact_vals"], order,
                                         method=O.NAIVE if method == "naive" else O.LIP, threads=8)
    n, s = (groups[0][3], groups[0][2]) if groups else (0, 0)
    return n, s, lines, rows


SCHEMAS = [
    # (n_fact, dim_rows, dim_sels, bloom_factor, htable_factor, gen)
    (300_000, [1000, 5000, 20_000], [499, 999, 9999], 2, 2, "reference"),      # 50 %, 20 %, 50 %
    ((1 << 20) + 3, [1 << 16, 3000, 777, 1 << 18], [1 << 14, 2999, 70, 1 << 16], 2, 4, "seeded"),
    (200_000, [4096], [40], 1, 2, "reference"),                                 # one dimension, 1 %
    (500_000, [100_000, 64], [90_000, 31], 4, 2, "seeded"),
]


@pytest.mark.parametrize("i", range(len(SCHEMAS)))
def test_star_schema_against_oracle(ctx, i):
    n_fact, rows, sels, bf, hf, gen = SCHEMAS[i]
    schema = StarSchema(n_fact, rows, sels, bloom_factor=bf, htable_factor=hf, gen=gen, seed=100 + i)
    _flush(f"star schema {i}: {n_fact} fact rows, dimensions {rows}, {gen}")
    cols = schema.host_columns()
    dev = StarSchema.to_device(cols)
    nd = schema.n_dims
    orders = [list(range(nd)), list(reversed(range(nd)))]
    for method in ("naive", "lip", "lip-adaptive"):
        for order in (orders if method != "lip-adaptive" else orders[:1]):
            r = ctx.star_join(schema.star_desc(dev, method=method, order=order))
            n, s, lines, rws = _oracle(schema, cols, "naive" if method == "naive" else "lip", order)
            assert (r["n_rows"], r["sum_total"]) == (n, s), (method, order)
            if method != "lip-adaptive":
                assert r["stats"]["lines_in"][:nd + 1] == lines[:nd + 1], (method, order)
                assert r["stats"]["rows_in"][:nd + 1] == rws[:nd + 1], (method, order)
    torch.cuda.synchronize()


def test_star_schema_batch_matches_alone(ctx):
    """Star schemas of different shapes as one batch (the batched mode's
    instances): each result equals the same query alone."""
    _flush("star schema batch")
    schemas = [StarSchema(50_000 * (k + 1), [257 * (k + 1), 4096, 61][: 1 + k % 3], [100, 2047, 30][: 1 + k % 3],
                          gen="seeded", seed=500 + k) for k in range(6)]
    devs = [StarSchema.to_device(s.host_columns()) for s in schemas]
    descs = [s.star_desc(d, method="lip") for s, d in zip(schemas, devs)]
    alone = [ctx.star_join(q) for q in descs]
    for graph in (False, True):
        b = ctx.batch(descs, use_graph=graph)
        for _ in range(2):
            res = b.run()
            for a, r in zip(alone, res):
                assert (a["n_rows"], a["sum_total"]) == (r["n_rows"], r["sum_total"])
        b.close()
    torch.cuda.synchronize()
