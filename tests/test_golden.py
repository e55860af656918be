"""Golden fixtures (tests/golden/*.json, written by make_fixtures.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP path
through the C-ABI must produce them (GPU)."""
import json
import os

import numpy as np
import pytest

import lipjoin as L
import oracle as O
from lipjoin import workloads as W
from q33_helpers import gpu_groups, oracle_q33, to_device

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(HERE, name + ".json")) as f:
        return json.load(f)


def _bloom(b):
    if b[0] == "mod":
        return ("mod", b[1]), O.bloom_spec(O.BLOOM_MOD, b[1], 1)
    return ("mult", b[1], b[2], b[3]), O.bloom_spec(O.BLOOM_MULT, b[1], b[2], layout=b[3])


def test_oracle_bloom_kat_fixture():
    fx = _load("bloom_kat")
    keys = np.arange(1024, dtype=np.int32)
    words = O.bloom_build(O.bloom_spec(O.BLOOM_MOD, 512, 1), keys, (keys % 3 == 0).astype(np.uint8))
    assert [int(w) for w in words] == fx["words"]
    assert fx["words"][:3] == [0xdb6db6db, 0xb6db6db6, 0x6db6db6d]


def test_oracle_q33_fixture():
    fx = _load("q33_small")
    host_dims = W.q33_host_dims(fx["seed"])
    host_fact = W.q33_host_fact(fx["seed"], fx["n_fact"])
    city = [tuple(s) for s in fx["city_pred"]]
    for name, method, order, bo in W.q33_plans():
        if name not in fx["plans"]:
            continue
        groups, lines, rows, oob = oracle_q33(host_dims, host_fact, order, method, bo, city_pred=city)
        assert [[g, list(p), s, c] for (g, p, s, c) in groups] == fx["groups"], name
        assert list(lines) == fx["plans"][name]["lines"] and list(rows) == fx["plans"][name]["rows"], name


def test_oracle_single_fixture():
    fx = _load("single_small")
    keys, attr, fk, meas = W.single_join_host(fx["seed"], fx["n_dim"], fx["n_fact"], 9999)
    for case in fx["cases"]:
        _, ob = _bloom(case["bloom"])
        r = O.single_join(O.DimSpec(keys, attr, fx["n_dim"], pred=list(W.single_join_pred(case["threshold"])),
                                    bloom=ob), fk, meas, method=O.LIP)
        assert {k: int(r[k]) for k in case["expect"]} == case["expect"]


@pytest.mark.gpu
def test_gpu_q33_fixture(ctx):
    fx = _load("q33_small")
    host_dims = W.q33_host_dims(fx["seed"])
    host_fact = W.q33_host_fact(fx["seed"], fx["n_fact"])
    dev = to_device({**host_dims, **host_fact})
    city = tuple(tuple(s) for s in fx["city_pred"])
    for name, method, order, bo in W.q33_plans():
        dims = W.q33_dim_descs(dev, city_pred=city)
        res = ctx.star_join(W.q33_star_desc(dev, fx["n_fact"], method, order, bo, dims=dims), groups_cap=1 << 14)
        assert [[g, list(p), s, c] for (g, p, s, c) in gpu_groups(res)] == fx["groups"], name
        if name in fx["plans"]:
            assert res["stats"]["lines_in"][:4] == fx["plans"][name]["lines"], name


@pytest.mark.gpu
def test_gpu_single_fixture(ctx):
    fx = _load("single_small")
    keys, attr, fk, meas = W.single_join_host(fx["seed"], fx["n_dim"], fx["n_fact"], 9999)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    for case in fx["cases"]:
        bl, _ = _bloom(case["bloom"])
        desc = L.dim_desc(dev["keys"], fx["n_dim"], fx["n_dim"], attr=dev["attr"],
                          predicate=W.single_join_pred(case["threshold"]), bloom=bl)
        r = ctx.hash_join(dev["fk"], dev["meas"], fx["n_fact"], desc, method="lip")
        assert {k: int(r[k]) for k in case["expect"]} == case["expect"]
