"""Shared builders for the Q3.3 / single-join parity tests (oracle side + GPU side)."""
import numpy as np

import oracle as O
from lipjoin import workloads as W

ORACLE_THREADS = 8


def oracle_q33_dims(host_dims, city_pred=None, year_pred=None, bloom=True):
    cp, yp = W.q33_predicates()
    city_pred = list(city_pred or cp)
    year_pred = list(year_pred or yp)
    d_len = W.Q33_DATE_MAX - W.Q33_DATE_MIN + 1
    bl = (lambda n, m=0: O.bloom_spec(O.BLOOM_MOD, O.round_up_pow2(n), 1, keys_min=m)) if bloom else \
        (lambda n, m=0: O.bloom_spec())
    return [
        O.DimSpec(host_dims["s_suppkey"], host_dims["s_city"], W.Q33_S_LEN, pred=city_pred, bloom=bl(W.Q33_S_LEN)),
        O.DimSpec(host_dims["c_custkey"], host_dims["c_city"], W.Q33_C_LEN, pred=city_pred, bloom=bl(W.Q33_C_LEN)),
        O.DimSpec(host_dims["d_datekey"], host_dims["d_year"], d_len, pred=year_pred, keys_min=W.Q33_DATE_MIN,
                  bloom=bl(d_len, W.Q33_DATE_MIN)),
    ]


def oracle_q33(host_dims, host_fact, order, method, bloom_order=None, city_pred=None, year_pred=None,
               threads=ORACLE_THREADS):
    dims = oracle_q33_dims(host_dims, city_pred, year_pred)
    fks = [host_fact["lo_suppkey"], host_fact["lo_custkey"], host_fact["lo_orderdate"]]
    return O.star_join(dims, fks, host_fact["lo_revenue"], order,
                       method=O.LIP if method != "naive" else O.NAIVE, bloom_order=bloom_order,
                       group_mult=W.Q33_GROUP_MULT, group_offset=W.Q33_GROUP_OFFSET, n_groups=W.Q33_N_GROUPS,
                       threads=threads)


def gpu_groups(res):
    """(g, payloads, sum, count) from Context.star_join, for comparison with the oracle."""
    return [(g, tuple(p), s, c) for (g, p, s, c, _) in res["groups"]]


def to_device(arrays, device="cuda"):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in arrays.items()}
