"""Oracle side of the SSB flight 2/3 parity tests, and an independent numpy
evaluation of each query written from its SQL-level meaning (used to pin the
oracle's predicate-step and group-index handling for these queries)."""
import numpy as np

import oracle as O
from lipjoin import ssb

SF = 1


def oracle_dims(q, host_dims, sf=SF):
    out = []
    for dq in q.dims:
        ht_len, kmin, bits = ssb.dim_params(dq, sf)
        out.append(O.DimSpec(host_dims[f"{dq.table}_key"], host_dims[dq.attr] if dq.attr else None, ht_len,
                             pred=list(dq.pred), payload=host_dims[dq.payload] if dq.payload else None,
                             keys_min=kmin, bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1, keys_min=kmin),
                             ht_kind=O.HT_PHT2 if dq.payload else O.HT_PHT1))
    return out


def oracle_query(q, host_dims, host_fact, order, method, sf=SF, threads=8):
    """(groups, lines, rows, n_oob) from the oracle's star join; a query with a
    subtracted measure runs twice and returns the per-group differences."""
    def one(measure):
        return O.star_join(oracle_dims(q, host_dims, sf), [host_fact[dq.fk] for dq in q.dims], measure,
                           order, method=O.NAIVE if method == "naive" else O.LIP, group_mult=q.group_mult,
                           group_offset=q.group_offset, n_groups=q.n_groups, threads=threads)
    groups, lines, rows, oob = one(host_fact["lo_revenue"])
    if q.measure_sub is not None:
        sub = one(host_fact[q.measure_sub])[0]
        assert [(a[0], a[3]) for a in groups] == [(b[0], b[3]) for b in sub]
        groups = [(a[0], a[1], a[2] - b[2], a[3]) for a, b in zip(groups, sub)]
    return groups, lines, rows, oob


def canon(q, groups):
    """Groups with the payloads of dimensions outside the group key (multiplier
    0) zeroed: such a payload is any selected row's value, not a result."""
    keep = [m != 0 for m in q.group_mult]
    return [(g[0], tuple(p if k else 0 for p, k in zip(g[1], keep))) + tuple(g[2:]) for g in groups]


def _row_of(table, key, host_dims):
    if table == "d":
        return np.searchsorted(host_dims["d_key"], key)
    return key.astype(np.int64) - 1           # keys 1..n in row order


def brute_force(name, host_dims, host_fact):
    """[(g, payloads, sum, count)] ascending g, from the query's SQL meaning:
    q21 s_region=1, p_category=1            group (d_year, p_brand1)
    q22 s_region=2, p_brand1 in [260,267]   group (d_year, p_brand1)
    q23 s_region=3, p_brand1=260            group (d_year, p_brand1)
    q31 c_region=s_region=2, d_year<=1997   group (c_nation, s_nation, d_year)
    q32 c_nation=s_nation=24, 1992<=d_year<=1997   group (c_city, s_city, d_year)
    q33 c_city, s_city in {231,235}, d_year<=1997  group (c_city, s_city, d_year)
    q34 c_city, s_city in {231,235}, d_yearmonthnum=199712
    q41 c_region=s_region=1, p_mfgr in {0,1}, SUM(revenue - supplycost)
                                            group (d_year, c_nation)
    q42 c_region=s_region=1, p_mfgr in {0,1}, d_year in {1997,1998}
                                            group (d_year, s_nation, p_category)
    q43 c_region=1, s_nation=24, p_category=3, d_year in {1997,1998}
                                            group (d_year, s_city, p_brand1)
    (d_year runs 1992..1998, so the reference's overwriting BlockPredLTE keeps
    1992..1997 as the SQL does)."""
    h, f = host_dims, host_fact
    s = _row_of("s", f["lo_suppkey"], h)
    d = _row_of("d", f["lo_orderdate"], h)
    year = h["d_year"][d].astype(np.int64)
    if name == "q41":
        c = _row_of("c", f["lo_custkey"], h)
        p = _row_of("p", f["lo_partkey"], h)
        mfgr, nation = h["p_mfgr"][p].astype(np.int64), h["c_nation"][c].astype(np.int64)
        region = h["s_region"][s].astype(np.int64)
        keep = (region == 1) & (h["c_region"][c] == 1) & ((mfgr == 0) | (mfgr == 1))
        g = (nation * 7 + year - 1992) % 175
        pay = np.stack([region, nation, mfgr, year], axis=1)
        measure = f["lo_revenue"].astype(np.int64) - f["lo_supplycost"].astype(np.int64)
        return _aggregate(g, pay, measure, keep)
    if name in ("q42", "q43"):
        c = _row_of("c", f["lo_custkey"], h)
        p = _row_of("p", f["lo_partkey"], h)
        c_region = h["c_region"][c].astype(np.int64)
        years = (year == 1997) | (year == 1998)
        measure = f["lo_revenue"].astype(np.int64) - f["lo_supplycost"].astype(np.int64)
        if name == "q42":
            nation, cat = h["s_nation"][s].astype(np.int64), h["p_category"][p].astype(np.int64)
            keep = (h["s_region"][s] == 1) & (c_region == 1) & (h["p_mfgr"][p] <= 1) & years
            g = ((year - 1992) * 625 + nation * 25 + cat) % 4375
            pay = np.stack([nation, c_region, cat, year], axis=1)
        else:
            city, brand = h["s_city"][s].astype(np.int64), h["p_brand1"][p].astype(np.int64)
            keep = (h["s_nation"][s] == 24) & (c_region == 1) & (h["p_category"][p] == 3) & years
            g = ((year - 1992) * 250000 + city * 1000 + brand) % 1750000
            pay = np.stack([city, c_region, brand, year], axis=1)
        return _aggregate(g, pay, measure, keep)
    if name.startswith("q2"):
        p = _row_of("p", f["lo_partkey"], h)
        brand, region = h["p_brand1"][p].astype(np.int64), h["s_region"][s]
        keep = {"q21": (region == 1) & (h["p_category"][p] == 1),
                "q22": (region == 2) & (brand >= 260) & (brand <= 267),
                "q23": (region == 3) & (brand == 260)}[name]
        g = (brand * 7 + year - 1992) % 7000
        pay = np.stack([region.astype(np.int64), brand, year], axis=1)
    else:
        c = _row_of("c", f["lo_custkey"], h)
        col, npay = ("nation", 25) if name == "q31" else ("city", 250)
        sv, cv = h[f"s_{col}"][s].astype(np.int64), h[f"c_{col}"][c].astype(np.int64)
        if name == "q31":
            keep = (h["s_region"][s] == 2) & (h["c_region"][c] == 2) & (year <= 1997)
        elif name == "q32":
            keep = (h["s_nation"][s] == 24) & (h["c_nation"][c] == 24) & (year >= 1992) & (year <= 1997)
        else:
            cities = lambda v: (v == 231) | (v == 235)
            date_ok = (year <= 1997) if name == "q33" else (h["d_yearmonthnum"][d] == 199712)
            keep = cities(sv) & cities(cv) & date_ok
        g = (sv * npay * 7 + cv * 7 + year - 1992) % (7 * npay * npay)
        pay = np.stack([sv, cv, year], axis=1)
    return _aggregate(g, pay, f["lo_revenue"].astype(np.int64), keep)


def _aggregate(g, pay, rev, keep):
    out = {}
    for gi, pi, ri in zip(g[keep], pay[keep], rev[keep]):
        e = out.setdefault(int(gi), [tuple(int(x) for x in pi), 0, 0])
        e[1] += int(ri)
        e[2] += 1
    return [(gi, e[0], e[1], e[2]) for gi, e in sorted(out.items())]
