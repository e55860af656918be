"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_fixtures.py
Each fixture holds the seeded inputs' parameters and the expected outputs;
the inputs themselves are regenerated from the seeds (lj_gen_host).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "crystal-opt_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from lipjoin import workloads as W  # noqa: E402
from q33_helpers import oracle_q33  # noqa: E402


def bloom_kat():
    """src/ops/bloom.cu:131-156 setup: keys 0..1023, every third key inserted, 16 words."""
    keys = np.arange(1024, dtype=np.int32)
    flags = (keys % 3 == 0).astype(np.uint8)
    spec = O.bloom_spec(O.BLOOM_MOD, 16 * 32, 1)
    words = O.bloom_build(spec, keys, flags)
    return {"keys": "0..1023", "inserted": "k % 3 == 0", "words": [int(w) for w in words]}


def q33_small():
    seed, n = W.seed_of(3), 1 << 16
    host_dims = W.q33_host_dims(seed)
    host_fact = W.q33_host_fact(seed, n)
    city = ((0, 0, 20),)       # city < 20: a few hundred groups at this size
    out = {"seed": seed, "n_fact": n, "city_pred": [list(s) for s in city], "groups": None, "plans": {}}
    for name, method, order, bo in W.q33_plans():
        if name == "lip-adaptive":
            continue
        groups, lines, rows, oob = oracle_q33(host_dims, host_fact, order, method, bo, city_pred=city)
        g = [[g, list(p), s, c] for (g, p, s, c) in groups]
        assert out["groups"] is None or out["groups"] == g, name   # every plan: the same groups
        out["groups"] = g
        out["plans"][name] = {"lines": list(lines), "rows": list(rows), "oob": oob}
    return out


def single_small():
    seed, nd, nf = W.seed_of(2), 1 << 12, 1 << 16
    keys, attr, fk, meas = W.single_join_host(seed, nd, nf, 9999)
    out = {"seed": seed, "n_dim": nd, "n_fact": nf, "cases": []}
    for thr, bloom in ((500, ("mult", 1 << 14, 2, 0)), (2000, ("mod", 1 << 12)), (100, ("mult", 1 << 13, 3, 1))):
        ob = O.bloom_spec(O.BLOOM_MOD, bloom[1], 1) if bloom[0] == "mod" else \
            O.bloom_spec(O.BLOOM_MULT, bloom[1], bloom[2], layout=bloom[3])
        pred = W.single_join_pred(thr)
        r = O.single_join(O.DimSpec(keys, attr, nd, pred=list(pred), bloom=ob), fk, meas, method=O.LIP)
        out["cases"].append({"threshold": thr, "bloom": list(bloom),
                             "expect": {k: int(r[k]) for k in ("n_selected", "n_bloom_pass", "n_match",
                                                                 "sum_measure", "sum_product")}})
    return out


def main():
    for name, fn in (("bloom_kat", bloom_kat), ("q33_small", q33_small), ("single_small", single_small)):
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fn(), f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()
