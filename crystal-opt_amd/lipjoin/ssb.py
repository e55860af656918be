"""The other SSB join-order queries (flights 2 and 3) as star-join descriptors.

Each query of src/ssb-jo/q{21,22,23,31,32,33,34}/{0..5}-{naive,lip}.cu is one
lj_star_desc: three dimensions, each with a predicate, a PHT_2 payload and a
HASH1 Bloom filter of roundUpToPowerOfTwo(ht_len) bits, and a dense group
index  g = (sum_d payload_d * group_mult[d] + group_offset) % n_groups.

Synthetic SSB-shaped tables at scale factor SF (src/ssb/ssb_utils.h:13-19
sizes: customer 30000*SF, supplier 2000*SF, part 200000*floor(1+log2 SF),
date 2556 calendar days), with the dictionary encodings the reference's
predicates assume:
  nation ~ U[0,24], region = nation // 5, city = nation*10 + U[0,9]
  category ~ U[0,24] (mfgr*5 + category), mfgr = category // 5,
  brand1 = category*40 + U[0,39]
  d_year = calendar year, d_yearmonthnum = datekey // 100
Lineorder: custkey ~ U[1,C], suppkey ~ U[1,S], partkey ~ U[1,P],
orderdate uniform over the 2556 datekeys, revenue ~ U[1,10^6],
supplycost ~ U[1,10^5].

The dimensions the reference builds as PHT_1 (keys only, probed by
BlockProbeAndPHT_1) are PHT_1 here too: the supplier in Q2.x (q21/0-lip.cu:77,
135), supplier and part in Q4.1 (q41/0-lip.cu:88,100,157,186), the customer
in Q4.2 (q42/0-naive.cu:62,128) and Q4.3 (q43/q43.cu:78,138).  None of them
enters the group key; their payload is reported as 0.  Q4.1 (q41/0-lip.cu) sums
(long long)(lo_revenue - lo_supplycost) (:127-128): the descriptor's
fact_measure_sub column is gathered with the measure and subtracted per row.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import _ptr
from . import (CMP_EQ, CMP_GTE, CMP_LTE, COMB_AND, COMB_OR, COMB_SET, GEN_DATEKEY, GEN_DATEKEY_SEQ,
               GEN_SEQUENCE, GEN_UNIFORM, GEN_YEAR_OF_DAY, HT_PHT1, HT_PHT2, dim_desc, gen_host,
               round_up_pow2, star_desc)

DATE_MIN, DATE_MAX = 19920101, 19981230
D_LEN = 2556
D_VAL_LEN = DATE_MAX - DATE_MIN + 1


def table_sizes(sf):
    """(customer, supplier, part, date) rows at scale factor sf; a 4-tuple of
    row counts (tables read from files) is returned as it is."""
    if isinstance(sf, tuple):
        return sf
    return 30000 * sf, 2000 * sf, 200000 * int(math.floor(1 + math.log2(sf))), D_LEN


def host_dims(seed, sf=1):
    """customer/supplier/part/date columns (numpy int32)."""
    c_len, s_len, p_len, d_len = table_sizes(sf)
    cols = {}
    for pre, n, st in (("c", c_len, 20), ("s", s_len, 23)):
        nation = gen_host(GEN_UNIFORM, seed, st, 0, 24, n)
        cols[f"{pre}_key"] = gen_host(GEN_SEQUENCE, seed, 0, 1, 0, n)
        cols[f"{pre}_nation"] = nation
        cols[f"{pre}_region"] = (nation // 5).astype(np.int32)
        cols[f"{pre}_city"] = (nation * 10 + gen_host(GEN_UNIFORM, seed, st + 1, 0, 9, n)).astype(np.int32)
    category = gen_host(GEN_UNIFORM, seed, 26, 0, 24, p_len)
    cols["p_key"] = gen_host(GEN_SEQUENCE, seed, 0, 1, 0, p_len)
    cols["p_category"] = category
    cols["p_mfgr"] = (category // 5).astype(np.int32)
    cols["p_brand1"] = (category * 40 + gen_host(GEN_UNIFORM, seed, 27, 0, 39, p_len)).astype(np.int32)
    cols["d_key"] = gen_host(GEN_DATEKEY_SEQ, seed, 0, 0, 0, d_len)
    cols["d_year"] = gen_host(GEN_YEAR_OF_DAY, seed, 0, 0, 0, d_len)
    cols["d_yearmonthnum"] = (cols["d_key"] // 100).astype(np.int32)
    return cols


def _fact_specs(sf):
    """lineorder column -> (generator kind, stream, lo, hi)."""
    c_len, s_len, p_len, d_len = table_sizes(sf)
    return {
        "lo_custkey": (GEN_UNIFORM, 30, 1, c_len),
        "lo_suppkey": (GEN_UNIFORM, 31, 1, s_len),
        "lo_partkey": (GEN_UNIFORM, 32, 1, p_len),
        "lo_orderdate": (GEN_DATEKEY, 33, 0, d_len),
        "lo_revenue": (GEN_UNIFORM, 34, 1, 1_000_000),
        "lo_supplycost": (GEN_UNIFORM, 35, 1, 100_000),
    }


def host_fact(seed, n, sf=1):
    """lineorder columns (numpy int32) of n rows."""
    return {k: gen_host(kind, seed, st, lo, hi, n) for k, (kind, st, lo, hi) in _fact_specs(sf).items()}


def device_columns(ctx, seed, n, sf=1, torch_device="cuda"):
    """Every table's columns as int32 device tensors: the dimensions generated on
    the host and copied, lineorder generated on the device (the same bits as
    host_fact)."""
    import torch
    cols = {k: torch.from_numpy(v).to(torch_device) for k, v in host_dims(seed, sf).items()}
    for k, (kind, st, lo, hi) in _fact_specs(sf).items():
        t = torch.empty(n, dtype=torch.int32, device=torch_device)
        ctx.gen_device(t, kind, seed, st, lo, hi)
        cols[k] = t
    return cols


# ------------------------------------------------------------------- queries --
@dataclass(frozen=True)
class DimQ:
    """One dimension of a query: its table, predicate column and steps, payload."""
    table: str          # "c", "s", "p" or "d"
    fk: str             # lineorder column joined to the table's key
    attr: str           # predicate column (None: every row selected)
    payload: str        # PHT_2 value column (None: a PHT_1 table, keys only)
    pred: tuple = ()    # ((comb, cmp, value), ...), crystal/pred.cuh semantics


@dataclass(frozen=True)
class Query:
    name: str
    dims: tuple                 # DimQ per descriptor dimension
    group_mult: tuple
    group_offset: int
    n_groups: int
    sum_bits: int               # width of the reference's atomicAdd: 32 (wraps) or 64
    letters: str                # dimension letters in descriptor order, for join-order names
    source: str                 # reference file:lines of the predicates and group index
    measure_sub: str = None     # lineorder column subtracted from lo_revenue (Q4.x)


_YEARS_SET = ((COMB_SET, CMP_GTE, 1992), (COMB_SET, CMP_LTE, 1997))   # BlockPredGTE, BlockPredLTE
_YEARS_AND = ((COMB_SET, CMP_GTE, 1992), (COMB_AND, CMP_LTE, 1997))   # BlockPredGTE, BlockPredAndLTE
_CITIES = ((COMB_SET, CMP_EQ, 231), (COMB_OR, CMP_EQ, 235))


def _flight2(name, s_region, part_pred, source):
    # descriptor dims (S, P, D); g = (brand*7 + (year-1992)) % (7*1000)
    return Query(name, (DimQ("s", "lo_suppkey", "s_region", None, ((COMB_SET, CMP_EQ, s_region),)),
                        DimQ("p", "lo_partkey", part_pred[0], "p_brand1", part_pred[1]),
                        DimQ("d", "lo_orderdate", None, "d_year", ())),
                 (0, 7, 1), -1992, 7 * 5 * 5 * 40, 64, "SPD", source)


def _flight3(name, s_pred, c_pred, d_pred, payload, n_pay, sum_bits, source):
    # descriptor dims (S, C, D); g = (s*n_pay*7 + c*7 + (year-1992)) % (7*n_pay*n_pay)
    s_attr, s_steps = s_pred
    c_attr, c_steps = c_pred
    d_attr, d_steps = d_pred
    return Query(name, (DimQ("s", "lo_suppkey", s_attr, f"s_{payload}", s_steps),
                        DimQ("c", "lo_custkey", c_attr, f"c_{payload}", c_steps),
                        DimQ("d", "lo_orderdate", d_attr, "d_year", d_steps)),
                 (n_pay * 7, 7, 1), -1992, 7 * n_pay * n_pay, sum_bits, "SCD", source)


QUERIES = {
    "q21": _flight2("q21", 1, ("p_category", ((COMB_SET, CMP_EQ, 1),)),
                    "q21/0-lip.cu:101-105,129,158,166,180-192"),
    "q22": _flight2("q22", 2, ("p_brand1", ((COMB_SET, CMP_GTE, 260), (COMB_AND, CMP_LTE, 267))),
                    "q22/0-lip.cu:101-105,129,157-160"),
    "q23": _flight2("q23", 3, ("p_brand1", ((COMB_SET, CMP_EQ, 260),)),
                    "q23/0-naive.cu:101-105,129,157"),
    "q31": _flight3("q31", ("s_region", ((COMB_SET, CMP_EQ, 2),)), ("c_region", ((COMB_SET, CMP_EQ, 2),)),
                    ("d_year", _YEARS_SET), "nation", 25, 64, "q31/0-lip.cu:102-109,135,166,196-199"),
    "q32": _flight3("q32", ("s_nation", ((COMB_SET, CMP_EQ, 24),)), ("c_nation", ((COMB_SET, CMP_EQ, 24),)),
                    ("d_year", _YEARS_AND), "city", 250, 32, "q32/0-lip.cu:102-108,133,164,194-197"),
    "q33": _flight3("q33", ("s_city", _CITIES), ("c_city", _CITIES), ("d_year", _YEARS_SET), "city", 250, 32,
                    "q33/0-lip.cu:98-111,132-135,192-195"),
    # descriptor dims (S, C, P, D), the order of q41/0-{naive,lip}.cu;
    # g = (c_nation*7 + (year-1992)) % (7*25)
    "q41": Query("q41", (DimQ("s", "lo_suppkey", "s_region", None, ((COMB_SET, CMP_EQ, 1),)),
                         DimQ("c", "lo_custkey", "c_region", "c_nation", ((COMB_SET, CMP_EQ, 1),)),
                         DimQ("p", "lo_partkey", "p_mfgr", None,
                              ((COMB_SET, CMP_EQ, 0), (COMB_OR, CMP_EQ, 1))),
                         DimQ("d", "lo_orderdate", None, "d_year", ())),
                 (0, 7, 0, 1), -1992, 7 * 25, 64, "SCPD",
                 "q41/0-lip.cu:120-128,151,157,178-186,209,217,237-243", measure_sub="lo_supplycost"),
    # q42/0-naive.cu: S->C->P->D; g = ((year-1992)*25*25 + s_nation*25 + category) % (7*25*25)
    "q42": Query("q42", (DimQ("s", "lo_suppkey", "s_region", "s_nation", ((COMB_SET, CMP_EQ, 1),)),
                         DimQ("c", "lo_custkey", "c_region", None, ((COMB_SET, CMP_EQ, 1),)),
                         DimQ("p", "lo_partkey", "p_mfgr", "p_category",
                              ((COMB_SET, CMP_EQ, 0), (COMB_OR, CMP_EQ, 1))),
                         DimQ("d", "lo_orderdate", "d_year", "d_year",
                              ((COMB_SET, CMP_EQ, 1997), (COMB_OR, CMP_EQ, 1998)))),
                 (25, 0, 1, 625), -1992 * 625, 7 * 25 * 25, 64, "SCPD",
                 "q42/0-naive.cu:94-101,123,150-152,181,210-213,254-268", measure_sub="lo_supplycost"),
    # q43/q43.cu: S->C->P->D; g = ((year-1992)*250*1000 + s_city*1000 + brand) % (7*250*1000),
    # int32 atomicAdd of revenue - supplycost
    "q43": Query("q43", (DimQ("s", "lo_suppkey", "s_nation", "s_city", ((COMB_SET, CMP_EQ, 24),)),
                         DimQ("c", "lo_custkey", "c_region", None, ((COMB_SET, CMP_EQ, 1),)),
                         DimQ("p", "lo_partkey", "p_category", "p_brand1", ((COMB_SET, CMP_EQ, 3),)),
                         DimQ("d", "lo_orderdate", "d_year", "d_year",
                              ((COMB_SET, CMP_EQ, 1997), (COMB_OR, CMP_EQ, 1998)))),
                 (1000, 0, 1, 250 * 1000), -1992 * 250 * 1000, 7 * 250 * 1000, 32, "SCPD",
                 "q43/q43.cu:105-111,133,160,189,218-221,262-276", measure_sub="lo_supplycost"),
    "q34": _flight3("q34", ("s_city", _CITIES), ("c_city", _CITIES),
                    ("d_yearmonthnum", ((COMB_SET, CMP_EQ, 199712),)), "city", 250, 32,
                    "q34/0-lip.cu:102-108,132-135,194-195"),
}


# The six join orders of q{21..34}/{0..5}-*.cu (header line 1), as permutations of
# the descriptor dimensions (X, Y, D): 0 X->Y->D, 1 X->D->Y, 2 D->X->Y, 3 D->Y->X,
# 4 Y->D->X, 5 Y->X->D, with X = supplier and Y = part (flight 2) or customer (flight 3).
ORDERS = {0: (0, 1, 2), 1: (0, 2, 1), 2: (2, 0, 1), 3: (2, 1, 0), 4: (1, 2, 0), 5: (1, 0, 2)}
# q41, q42 and q43 ship one order, S->C->P->D (q41/0-{naive,lip}.cu line 1,
# q42/0-naive.cu line 1, the probe of q43/q43.cu:70-93)
ORDERS_4 = {0: (0, 1, 2, 3)}


def orders_of(q):
    return ORDERS if len(q.dims) == 3 else ORDERS_4


def _ht_len(table, sf):
    c_len, s_len, p_len, _ = table_sizes(sf)
    return {"c": c_len, "s": s_len, "p": p_len, "d": D_VAL_LEN}[table]


def dim_params(dq, sf):
    """(ht_len, keys_min, bloom_bits) of one dimension, as runQuery sizes them
    (PHT_2 slots = table rows, the date table 19981230-19920101+1 slots with
    keys_min 19920101; Bloom bits = roundUpToPowerOfTwo(slots), e.g.
    q21/0-lip.cu:212-223)."""
    n = _ht_len(dq.table, sf)
    return n, (DATE_MIN if dq.table == "d" else 0), round_up_pow2(n)


def dim_descs(q, cols, sf, bloom=True):
    """lj_dim_desc per dimension from a dict of device columns (host_dims names)."""
    out = []
    for dq in q.dims:
        ht_len, kmin, bits = dim_params(dq, sf)
        n_rows = table_sizes(sf)[3] if dq.table == "d" else _ht_len(dq.table, sf)
        out.append(dim_desc(cols[f"{dq.table}_key"], n_rows, ht_len,
                            attr=cols[dq.attr] if dq.attr else None,
                            payload=cols[dq.payload] if dq.payload else None,
                            predicate=dq.pred, keys_min=kmin, bloom=("mod", bits) if bloom else None,
                            ht_kind=HT_PHT2 if dq.payload else HT_PHT1))
    return out


def star(q, cols, n_fact, method, order, sf, bloom_order=None, dims=None, resident=None):
    """lj_star_desc of query q over device columns (dimension and lineorder);
    resident: built dimensions (lj_dim_build) used instead of building in the call."""
    if dims is None and resident is None:
        dims = dim_descs(q, cols, sf)
    return star_desc([cols[dq.fk] for dq in q.dims], cols["lo_revenue"], n_fact, dims, order,
                     method=method, bloom_order=bloom_order, group_mult=q.group_mult,
                     group_offset=q.group_offset, n_groups=q.n_groups, resident=resident)


def run(ctx, q, cols, n_fact, method, order, sf, dims=None, resident=None, groups_cap=1 << 14):
    """lj_star_join of query q (Context.star_join's result dict).  A query with
    a subtracted measure (Q4.1) passes it as fact_measure_sub: the probe adds
    (int32)(lo_revenue - lo_supplycost) per row, in one pass."""
    desc = star(q, cols, n_fact, method, order, sf, dims=dims, resident=resident)
    if q.measure_sub is not None:
        desc.fact_measure_sub = _ptr(cols[q.measure_sub])
    return ctx.star_join(desc, groups_cap=groups_cap)


# ------------------------------------------------------------------ the CLI --
# Column files of the reference's loader (src/ssb/ssb_utils.h:106-153: table
# name + column index, raw little-endian int32), by the names used here.
COLUMN_FILES = {
    "lo_custkey": "LINEORDER2", "lo_partkey": "LINEORDER3", "lo_suppkey": "LINEORDER4",
    "lo_orderdate": "LINEORDER5", "lo_revenue": "LINEORDER12", "lo_supplycost": "LINEORDER13",
    "p_key": "PART0", "p_mfgr": "PART2", "p_category": "PART3", "p_brand1": "PART4",
    "s_key": "SUPPLIER0", "s_city": "SUPPLIER3", "s_nation": "SUPPLIER4", "s_region": "SUPPLIER5",
    "c_key": "CUSTOMER0", "c_city": "CUSTOMER3", "c_nation": "CUSTOMER4", "c_region": "CUSTOMER5",
    "d_key": "DDATE0", "d_year": "DDATE4", "d_yearmonthnum": "DDATE5",
}

# payload columns in the order the reference prints a result row (then the sum):
# q21-q23 year brand (q21/0-lip.cu:281-283), q31-q34 year c_x s_x,
# q41 year c_nation, q42 year s_nation category, q43 year s_city brand
PRINT_ORDER = {"q21": (2, 1), "q22": (2, 1), "q23": (2, 1), "q31": (2, 1, 0), "q32": (2, 1, 0),
               "q33": (2, 1, 0), "q34": (2, 1, 0), "q41": (3, 1), "q42": (3, 0, 2), "q43": (3, 0, 2)}


def read_columns(data_dir, names):
    """numpy int32 columns from the reference's column files; the row count of
    each table is the file's size / 4."""
    import os
    out = {}
    for k in names:
        path = os.path.join(data_dir, COLUMN_FILES[k])
        out[k] = np.fromfile(path, dtype="<i4")
    return out


def parse_plan(q, plan):
    """(method, order) of a plan name: <i>-naive / <i>-lip (the reference's
    binaries) or lip-adaptive."""
    orders = orders_of(q)
    if plan == "lip-adaptive":
        return "lip-adaptive", orders[0]
    i, _, m = plan.partition("-")
    if not i.isdigit() or int(i) not in orders or m not in ("naive", "lip"):
        raise ValueError(f"unknown plan {plan!r} for {q.name}")
    return m, orders[int(i)]


def main(argv=None):
    import argparse
    import time
    ap = argparse.ArgumentParser(
        prog="python -m lipjoin.ssb",
        description="SSB join-order query over the C-ABI; a stand-in for the reference's "
                    "bin/ssb-jo/<query>/<i>-<method> binaries (same --t, Result rows and "
                    '{"query":N,"time_query":ms} lines).')
    ap.add_argument("--query", required=True, choices=sorted(QUERIES))
    ap.add_argument("--plan", default="0-lip", help="<i>-naive | <i>-lip | lip-adaptive")
    ap.add_argument("--t", type=int, default=3, help="trials")
    ap.add_argument("--data_dir", help="directory of the reference's column files")
    ap.add_argument("--sf", type=int, default=20, help="scale factor of the synthetic tables")
    ap.add_argument("--n", type=int, default=120_000_000, help="synthetic lineorder rows")
    ap.add_argument("--i32", action="store_true", help="print the int32 (wrapping) sums")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)

    import torch
    from . import Context
    from . import workloads as W
    q = QUERIES[args.query]
    method, order = parse_plan(q, args.plan)
    ctx = Context(args.device)
    if args.data_dir:
        need = {dq.fk for dq in q.dims} | {f"{dq.table}_key" for dq in q.dims} | {"lo_revenue"}
        need |= {dq.attr for dq in q.dims if dq.attr} | {dq.payload for dq in q.dims if dq.payload}
        if q.measure_sub:
            need.add(q.measure_sub)
        host = read_columns(args.data_dir, sorted(need))
        sizes = tuple(host[f"{t}_key"].size if f"{t}_key" in host else 0 for t in "cspd")
        cols = {k: torch.from_numpy(v).to(f"cuda:{args.device}") for k, v in host.items()}
        n, sf = host["lo_revenue"].size, sizes
        print("** LOADED DATA **")
    else:
        cols = device_columns(ctx, W.seed_of(7), args.n, args.sf, f"cuda:{args.device}")
        n, sf = args.n, args.sf
        print("** GENERATED DATA **")
    torch.cuda.synchronize()
    print("** LOADED DATA TO GPU **")
    for _ in range(args.t):
        t0 = time.perf_counter()
        r = run(ctx, q, cols, n, method, order, sf, groups_cap=1 << 20)
        wall = (time.perf_counter() - t0) * 1e3
        print("Result:")
        for g, pay, s, _c, s32 in r["groups"]:
            print(" ".join(str(pay[i]) for i in PRINT_ORDER[q.name]), s32 if args.i32 else s)
        print(f"Res Count: {r['n_groups']}")
        print(f"Time Taken Total: {wall:g}")
        print(f'{{"query":{q.name[1:]},"time_query":{r["time_query_ms"]:g}}}')
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
