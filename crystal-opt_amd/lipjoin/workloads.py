"""Seeded synthetic workloads of the star-join path and the SSB Q3.3 query.

Column streams are counter-based (lj_gen_*): the host and the device produce
the same bits, so the CPU oracle and the GPU see identical inputs.

  single_join  : dim keys 0..D-1 in shuffled row order, attr ~ U[0, attr_hi],
                 fact fk ~ U[0, D-1], measure ~ U[1, 10000].
                 (2^20 dim, 2^24 fact, attr<10 over U[0,99]: the plain join;
                  2^22 dim, 2^28 fact, attr<100*s over U[0,9999]: the Bloom join)
  q33          : lineorder(custkey, suppkey, orderdate, partkey, revenue) over
                 customer(custkey 1..600000, city ~ U[0,249]),
                 supplier(suppkey 1..40000, city ~ U[0,249]),
                 date(2556 calendar days from 1992-01-01, year).
The Q3.3 query is test/ssb/queries/transformed/q33.sql:1-9 as the reference's
drivers run it (src/ssb-jo/q33/0-lip.cu:114-203, 205-302):
  s_city = 231 or 235 ; c_city = 231 or 235 ; d_year >= 1992 then (overwrite)
  d_year <= 1997 ; group (s_city*250*7 + c_city*7 + (year-1992)) % (7*250*250),
  PHT_2 slots: s_len, c_len, 19981230-19920101+1 with keys_min 19920101;
  Bloom bits roundUpToPowerOfTwo of the same lengths (HASH1, k = 1).
"""
import numpy as np

from . import (BLOOM_MOD, CMP_EQ, CMP_GTE, CMP_LTE, CMP_LT, COMB_OR, COMB_SET, GEN_DATEKEY,
               GEN_DATEKEY_SEQ, GEN_SEQUENCE, GEN_UNIFORM, GEN_YEAR_OF_DAY, dim_desc, gen_host,
               round_up_pow2, shuffle_host, star_desc)

SEED_BASE = 0x1A7E5EED00


def seed_of(config):
    return SEED_BASE + int(config)


# ---------------------------------------------------------------- column specs --
# name -> (kind, stream, lo, hi)
def single_join_specs(n_dim, attr_hi):
    return {
        "dim_attr": (GEN_UNIFORM, 1, 0, attr_hi),
        "fact_fk": (GEN_UNIFORM, 2, 0, n_dim - 1),
        "fact_measure": (GEN_UNIFORM, 3, 1, 10000),
    }


Q33_C_LEN, Q33_S_LEN, Q33_D_LEN = 600000, 40000, 2556
Q33_DATE_MIN, Q33_DATE_MAX = 19920101, 19981230
Q33_FACT_ROWS = 120_000_000

Q33_FACT_SPECS = {
    "lo_custkey": (GEN_UNIFORM, 10, 1, Q33_C_LEN),
    "lo_suppkey": (GEN_UNIFORM, 11, 1, Q33_S_LEN),
    "lo_orderdate": (GEN_DATEKEY, 12, 0, Q33_D_LEN),
    "lo_partkey": (GEN_UNIFORM, 13, 1, 1_000_000),
    "lo_revenue": (GEN_UNIFORM, 14, 1, 1_000_000),
}
Q33_DIM_SPECS = {
    "c_custkey": (GEN_SEQUENCE, 0, 1, 0, Q33_C_LEN),
    "c_city": (GEN_UNIFORM, 1, 0, 249, Q33_C_LEN),
    "s_suppkey": (GEN_SEQUENCE, 0, 1, 0, Q33_S_LEN),
    "s_city": (GEN_UNIFORM, 2, 0, 249, Q33_S_LEN),
    "d_datekey": (GEN_DATEKEY_SEQ, 0, 0, 0, Q33_D_LEN),
    "d_year": (GEN_YEAR_OF_DAY, 0, 0, 0, Q33_D_LEN),
}

# dimension indices of the star descriptor
S, C_, D = 0, 1, 2
# join orders of src/ssb-jo/q33/{0..5}-*.cu (header line 1 of each file)
Q33_ORDERS = {
    0: (S, C_, D),  # S -> C -> D
    1: (S, D, C_),  # S -> D -> C
    2: (D, S, C_),  # D -> S -> C
    3: (D, C_, S),  # D -> C -> S
    4: (C_, D, S),  # C -> D -> S
    5: (C_, S, D),  # C -> S -> D
}
ORDER_NAMES = {0: "SCD", 1: "SDC", 2: "DSC", 3: "DCS", 4: "CDS", 5: "CSD"}
# 3-lip-best-bloom.cu: Bloom probes S -> C -> D, hash probes D -> C -> S
Q33_BEST_BLOOM = {"bloom_order": (S, C_, D), "order": (D, C_, S)}
Q33_GROUP_MULT = (250 * 7, 7, 1)   # s_city, c_city, year
Q33_GROUP_OFFSET = -1992
Q33_N_GROUPS = (1998 - 1992 + 1) * 250 * 250


def q33_host_dims(seed):
    cols = {}
    for name, (kind, stream, lo, hi, n) in Q33_DIM_SPECS.items():
        cols[name] = gen_host(kind, seed, stream, lo, hi, n)
    return cols


def q33_host_fact(seed, n, names=None):
    names = names or list(Q33_FACT_SPECS)
    return {k: gen_host(*Q33_FACT_SPECS[k][:1], seed, *Q33_FACT_SPECS[k][1:], n) for k in names}


def q33_device_columns(ctx, seed, n_fact, torch_device="cuda"):
    """All Q3.3 columns generated on the device (int32 torch tensors)."""
    import torch
    cols = {}
    for name, (kind, stream, lo, hi, n) in Q33_DIM_SPECS.items():
        t = torch.empty(n, dtype=torch.int32, device=torch_device)
        ctx.gen_device(t, kind, seed, stream, lo, hi)
        cols[name] = t
    for name, (kind, stream, lo, hi) in Q33_FACT_SPECS.items():
        t = torch.empty(n_fact, dtype=torch.int32, device=torch_device)
        ctx.gen_device(t, kind, seed, stream, lo, hi)
        cols[name] = t
    return cols


def q33_predicates(cities=(231, 235), years=(1992, 1997)):
    city = ((COMB_SET, CMP_EQ, cities[0]), (COMB_OR, CMP_EQ, cities[1]))
    # build_hashtable_d: BlockPredGTE then BlockPredLTE, which overwrites
    year = ((COMB_SET, CMP_GTE, years[0]), (COMB_SET, CMP_LTE, years[1]))
    return city, year


def q33_dim_descs(cols, city_pred=None, year_pred=None, bloom=True):
    """(S, C, D) dimension descriptors from a dict of device (or host-address) columns."""
    cp, yp = q33_predicates()
    city_pred = city_pred or cp
    year_pred = year_pred or yp
    d_val_len = Q33_DATE_MAX - Q33_DATE_MIN + 1
    b = (lambda n: ("mod", round_up_pow2(n))) if bloom else (lambda n: None)
    s = dim_desc(cols["s_suppkey"], Q33_S_LEN, Q33_S_LEN, attr=cols["s_city"], predicate=city_pred,
                 bloom=b(Q33_S_LEN))
    c = dim_desc(cols["c_custkey"], Q33_C_LEN, Q33_C_LEN, attr=cols["c_city"], predicate=city_pred,
                 bloom=b(Q33_C_LEN))
    d = dim_desc(cols["d_datekey"], Q33_D_LEN, d_val_len, attr=cols["d_year"], predicate=year_pred,
                 keys_min=Q33_DATE_MIN, bloom=b(d_val_len))
    return [s, c, d]


def q33_star_desc(cols, n_fact, method, order, bloom_order=None, dims=None, resident=None):
    dims = dims if dims is not None or resident is not None else q33_dim_descs(cols)
    return star_desc([cols["lo_suppkey"], cols["lo_custkey"], cols["lo_orderdate"]], cols["lo_revenue"],
                     n_fact, dims, order, method=method, bloom_order=bloom_order,
                     group_mult=Q33_GROUP_MULT, group_offset=Q33_GROUP_OFFSET, n_groups=Q33_N_GROUPS,
                     resident=resident)


def q33_plans():
    """The 13 reference plans: 6 orders x {naive, lip} + best-bloom, plus the adaptive LIP."""
    plans = []
    for i, o in Q33_ORDERS.items():
        plans.append((f"{i}-naive", "naive", o, None))
        plans.append((f"{i}-lip", "lip", o, None))
    plans.append(("3-lip-best-bloom", "lip", Q33_BEST_BLOOM["order"], Q33_BEST_BLOOM["bloom_order"]))
    plans.append(("lip-adaptive", "lip-adaptive", Q33_ORDERS[0], None))
    return plans


def single_join_host(seed, n_dim, n_fact, attr_hi):
    keys = np.arange(n_dim, dtype=np.int32)
    shuffle_host(keys, seed, 0)
    sp = single_join_specs(n_dim, attr_hi)
    attr = gen_host(sp["dim_attr"][0], seed, *sp["dim_attr"][1:], n_dim)
    fk = gen_host(sp["fact_fk"][0], seed, *sp["fact_fk"][1:], n_fact)
    meas = gen_host(sp["fact_measure"][0], seed, *sp["fact_measure"][1:], n_fact)
    return keys, attr, fk, meas


def single_join_device(ctx, seed, n_dim, n_fact, attr_hi, torch_device="cuda"):
    """Dimension generated on the host (shuffle), fact columns on the device."""
    import torch
    keys = np.arange(n_dim, dtype=np.int32)
    shuffle_host(keys, seed, 0)
    sp = single_join_specs(n_dim, attr_hi)
    attr = gen_host(sp["dim_attr"][0], seed, *sp["dim_attr"][1:], n_dim)
    d_keys = torch.from_numpy(keys).to(torch_device)
    d_attr = torch.from_numpy(attr).to(torch_device)
    fk = torch.empty(n_fact, dtype=torch.int32, device=torch_device)
    meas = torch.empty(n_fact, dtype=torch.int32, device=torch_device)
    ctx.gen_device(fk, sp["fact_fk"][0], seed, *sp["fact_fk"][1:])
    ctx.gen_device(meas, sp["fact_measure"][0], seed, *sp["fact_measure"][1:])
    return d_keys, d_attr, fk, meas


def single_join_pred(threshold):
    """attr < threshold (the configs' predicate; the reference uses <= num_select)."""
    return ((COMB_SET, CMP_LT, int(threshold)),)


# ------------------------------------------------------------------ batches --
# Config 4: many independent single joins shaped like the Bloom join: fact
# rows floor(2^(lo + span*u)), u ~ U[0,1); a dimension of n_dim dense keys per
# instance (attr ~ U[0,9999]); selectivity s ~ U[s_lo, s_hi] percent with the
# predicate attr < round(100*s); the Bloom filter chosen by the autotuner.
def batch_specs(seed, n_inst=64, log_lo=20, log_span=7, n_dim=1 << 22, s_lo=0.5, s_hi=20.0):
    rng = np.random.default_rng(seed)
    u = rng.random(n_inst)
    sel = rng.uniform(s_lo, s_hi, n_inst)
    out = []
    for i in range(n_inst):
        out.append({"n_fact": int(np.floor(2.0 ** (log_lo + log_span * u[i]))), "n_dim": int(n_dim),
                    "threshold": int(round(100 * sel[i])), "sel_pct": float(sel[i]),
                    "seed": int(seed) * 1000 + i})
    return out


def batch_instance_host(spec):
    """Host columns of one instance (the same bits as batch_instance_device)."""
    n_dim, n_fact, seed = spec["n_dim"], spec["n_fact"], spec["seed"]
    keys = np.arange(n_dim, dtype=np.int32)
    sp = single_join_specs(n_dim, 9999)
    attr = gen_host(sp["dim_attr"][0], seed, *sp["dim_attr"][1:], n_dim)
    fk = gen_host(sp["fact_fk"][0], seed, *sp["fact_fk"][1:], n_fact)
    meas = gen_host(sp["fact_measure"][0], seed, *sp["fact_measure"][1:], n_fact)
    return keys, attr, fk, meas


def batch_instance_device(ctx, spec, torch_device="cuda"):
    import torch
    n_dim, n_fact, seed = spec["n_dim"], spec["n_fact"], spec["seed"]
    sp = single_join_specs(n_dim, 9999)
    cols = {}
    cols["keys"] = torch.empty(n_dim, dtype=torch.int32, device=torch_device)
    ctx.gen_device(cols["keys"], GEN_SEQUENCE, seed, 0, 0, 0)
    for name, key, n in (("attr", "dim_attr", n_dim), ("fk", "fact_fk", n_fact), ("meas", "fact_measure", n_fact)):
        cols[name] = torch.empty(max(n, 1), dtype=torch.int32, device=torch_device)
        ctx.gen_device(cols[name], sp[key][0], seed, *sp[key][1:], n=n)
    return cols


def batch_desc(cols, spec, method="lip", bloom=None):
    """Star descriptor (one dimension, one group) of a batch instance."""
    d = dim_desc(cols["keys"], spec["n_dim"], spec["n_dim"], attr=cols["attr"],
                 predicate=single_join_pred(spec["threshold"]), bloom=bloom)
    return star_desc([cols["fk"]], cols["meas"], spec["n_fact"], [d], (0,), method=method)


def tuned_bloom(ctx, spec, semi_join=False):
    """The autotuner's Bloom filter for an instance, as a dim_desc bloom tuple.
    semi_join: the join needs no payload (one group, no product)."""
    n_sel = int(round(spec["n_dim"] * spec["threshold"] / 10000.0))
    ch = ctx.tune_bloom(n_sel, spec["n_dim"], spec["threshold"] / 10000.0, semi_join=semi_join)
    if ch["bloom_hash"] == BLOOM_MOD:
        return ("mod", ch["bloom_bits"]), ch
    return ("mult", ch["bloom_bits"], ch["bloom_k"], ch["bloom_layout"]), ch


# ------------------------------------------------------------------ serving --
# Config 5: requests over resident dimensions.  A seeded pool of n_templates
# predicate templates (attr < round(100*s), s ~ U[0.5, 20] %) over one
# dimension of n_dim dense keys, each built once (PHT_2 + the autotuner's
# Bloom filter); a resident fact pool of n_pool rows; a request is a template
# and a slice [offset, offset + rows) of the pool, rows = floor(2^(lo + span*u)).
def serve_templates(ctx, seed, n_dim=1 << 22, n_templates=8, torch_device="cuda"):
    spec = {"n_dim": n_dim, "n_fact": 0, "seed": int(seed), "threshold": 0}
    cols = batch_instance_device(ctx, spec, torch_device)
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n_templates):
        thr = int(round(100 * rng.uniform(0.5, 20.0)))
        sp = dict(spec, threshold=thr)
        bloom, _ = tuned_bloom(ctx, sp, semi_join=True)    # requests sum one group
        d = dim_desc(cols["keys"], n_dim, n_dim, attr=cols["attr"], predicate=single_join_pred(thr), bloom=bloom)
        out.append({"dim": ctx.build_dim(d), "threshold": thr, "bloom": bloom})
    return cols, out


def serve_pool(ctx, seed, n_pool, n_dim=1 << 22, torch_device="cuda"):
    spec = {"n_dim": n_dim, "n_fact": n_pool, "seed": int(seed) + 1, "threshold": 0}
    return batch_instance_device(ctx, spec, torch_device)


def serve_requests(seed, n_req, n_templates, n_pool, log_lo=18, log_span=8):
    """[(template, offset, rows)]: offsets are multiples of 4 (16-byte aligned slices)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_req):
        rows = int(np.floor(2.0 ** (log_lo + log_span * rng.random())))
        rows = min(rows, n_pool)
        off = int(rng.integers(0, (n_pool - rows) // 4 + 1)) * 4
        out.append((int(rng.integers(0, n_templates)), off, rows))
    return out


def serve_desc(pool, tmpl, offset, rows):
    fk = pool["fk"].data_ptr() + 4 * offset
    meas = pool["meas"].data_ptr() + 4 * offset
    return star_desc([fk], meas, rows, None, (0,), method="lip", resident=[tmpl["dim"]])


__all__ = [n for n in dir() if not n.startswith("_")] + ["BLOOM_MOD"]
