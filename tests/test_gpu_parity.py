"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on the
same seeded inputs.  Integer work, so every comparison is exact: Bloom words,
PHT words, survivor counts, per-stage row/line statistics, int64 group sums
and group keys.  Sizes here let the oracle finish in seconds; the full-size
runs are in test_gpu_fullsize.py."""
import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O
from lipjoin import workloads as W
from q33_helpers import gpu_groups, oracle_q33, to_device

pytestmark = pytest.mark.gpu

SEED = W.seed_of(3)


def _flush(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ generators --
@pytest.mark.parametrize("spec", [(L.GEN_UNIFORM, 3, 1, 600000), (L.GEN_UNIFORM, 4, -2**31, 2**31 - 1),
                                  (L.GEN_DATEKEY, 12, 0, 2556), (L.GEN_SEQUENCE, 0, 1, 0),
                                  (L.GEN_YEAR_OF_DAY, 0, 0, 0), (L.GEN_DATEKEY_SEQ, 0, 0, 0)])
def test_gen_device_matches_host(ctx, spec):
    kind, stream, lo, hi = spec
    n = 100003
    t = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_device(t, kind, 77, stream, lo, hi)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), L.gen_host(kind, 77, stream, lo, hi, n))


# ------------------------------------------------------------ dimension builds --
BLOOMS = [("mod", 65536), ("mod", 1000), ("mult", 1 << 12, 1, L.BLOOM_STANDARD),
          ("mult", 1 << 14, 2, L.BLOOM_STANDARD), ("mult", 1 << 16, 3, L.BLOOM_STANDARD),
          ("mult", 1 << 12, 1, L.BLOOM_BLOCKED), ("mult", 1 << 13, 2, L.BLOOM_BLOCKED),
          ("mult", 1 << 15, 3, L.BLOOM_BLOCKED)]


def _oracle_bloom(b, keys_min=0):
    if b[0] == "mod":
        return O.bloom_spec(O.BLOOM_MOD, b[1], 1, keys_min=keys_min)
    return O.bloom_spec(O.BLOOM_MULT, b[1], b[2], layout=b[3], keys_min=keys_min)


@pytest.mark.parametrize("bloom", BLOOMS)
def test_dim_build_tables_bit_exact(ctx, bloom):
    _flush(f"dim build {bloom}")
    n = 50000
    keys = np.arange(5, n + 5, dtype=np.int32)
    L.shuffle_host(keys, 1, 0)
    attr = L.gen_host(L.GEN_UNIFORM, 1, 1, 0, 99, n)
    dk, da = torch.from_numpy(keys).cuda(), torch.from_numpy(attr).cuda()
    pred = ((L.COMB_SET, L.CMP_GTE, 20), (L.COMB_AND, L.CMP_LT, 45))
    desc = L.dim_desc(dk, n, 40000, attr=da, predicate=pred, keys_min=3, bloom=bloom)
    dim = ctx.build_dim(desc)
    try:
        ht, bf = dim.export()
        info = dim.info()
    finally:
        dim.release()
    flags = O.predicate(attr, list(pred))
    ht_ref, oob = O.pht2_build(keys, attr, flags, 40000, 3)
    # ht_len 40000 < key range: slots collide, but no two selected keys hit the
    # same slot unless equal mod 40000 -- compare only key words (CAS keeps the first
    # writer, whose order differs); values are compared where the slot is unshared.
    sel = keys[flags.astype(bool)]
    slots = (sel.astype(np.int64) - 3) % 40000
    uniq, cnt = np.unique(slots, return_counts=True)
    single = uniq[cnt == 1]
    assert np.array_equal(ht[2 * single], ht_ref[2 * single])
    assert np.array_equal(ht[2 * single + 1], ht_ref[2 * single + 1])
    empty = np.setdiff1d(np.arange(40000), uniq)
    assert not ht[2 * empty].any() and not ht[2 * empty + 1].any()
    bf_ref = O.bloom_build(_oracle_bloom(bloom, 3), keys, flags)
    assert np.array_equal(bf, bf_ref)
    assert info["n_selected"] == int(flags.sum())
    assert info["bloom_bits_set"] == sum(bin(int(w)).count("1") for w in bf_ref)


def test_dim_build_perfect_hash_exact(ctx):
    """Dense keys 1..n into n slots: a perfect hash, so the whole PHT is exact."""
    n = 600000
    keys = np.arange(1, n + 1, dtype=np.int32)
    city = L.gen_host(L.GEN_UNIFORM, SEED, 1, 0, 249, n)
    dk, dc = torch.from_numpy(keys).cuda(), torch.from_numpy(city).cuda()
    cp, _ = W.q33_predicates()
    dim = ctx.build_dim(L.dim_desc(dk, n, n, attr=dc, predicate=cp, bloom=("mod", L.round_up_pow2(n))))
    ht, bf = dim.export()
    dim.release()
    flags = O.predicate(city, list(cp))
    ht_ref, _ = O.pht2_build(keys, city, flags, n, 0)
    assert np.array_equal(ht, ht_ref)
    assert np.array_equal(bf, O.bloom_build(O.bloom_spec(O.BLOOM_MOD, 1 << 20, 1), keys, flags))


# ------------------------------------------------------------------ single join --
def _single_case(ctx, n_dim, n_fact, attr_hi, thr, bloom, method, seed=SEED):
    keys, attr, fk, meas = W.single_join_host(seed, n_dim, n_fact, attr_hi)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    pred = W.single_join_pred(thr)
    desc = L.dim_desc(dev["keys"], n_dim, n_dim, attr=dev["attr"], predicate=pred, bloom=bloom)
    r = ctx.hash_join(dev["fk"], dev["meas"], n_fact, desc, method=method)
    ob = _oracle_bloom(bloom) if bloom else O.bloom_spec()
    ref = O.single_join(O.DimSpec(keys, attr, n_dim, pred=list(pred), bloom=ob), fk, meas,
                        method=O.NAIVE if method == "naive" else O.LIP, threads=8)
    return r, ref


@pytest.mark.parametrize("method", ["naive", "lip", "lip-adaptive"])
@pytest.mark.parametrize("bloom", [None, ("mod", 1 << 14), ("mult", 1 << 16, 2, L.BLOOM_STANDARD),
                                   ("mult", 1 << 18, 3, L.BLOOM_BLOCKED), ("mult", 1 << 22, 1, L.BLOOM_STANDARD)])
def test_single_join_parity(ctx, method, bloom):
    _flush(f"single join {method} {bloom}")
    r, ref = _single_case(ctx, 1 << 16, (1 << 20) + 13, 99, 10, bloom, method)
    assert r["n_selected"] == ref["n_selected"]
    assert (r["n_match"], r["sum_measure"], r["sum_product"]) == \
        (ref["n_match"], ref["sum_measure"], ref["sum_product"])
    if method != "naive":
        assert r["n_bloom_pass"] == ref["n_bloom_pass"]
    if method != "lip-adaptive":
        st = r["stats"]
        assert st["rows_in"][0] == (1 << 20) + 13
        assert st["lines_in"][0] == ref["lines_fk"]
        assert st["lines_in"][1] == ref["lines_measure"]
        assert st["rows_in"][1] == ref["n_match"]


def test_single_join_key_zero_quirk(ctx):
    """join.cuh:123-128 -- the dimension row (key 0, attr 0) can never match."""
    n = 1024
    keys = np.arange(n, dtype=np.int32)
    attr = np.zeros(n, np.int32)
    attr[1:] = 5
    fk = np.tile(keys, 8).astype(np.int32)
    meas = np.ones(fk.size, np.int32)
    dev = to_device({"keys": keys, "attr": attr, "fk": fk, "meas": meas})
    desc = L.dim_desc(dev["keys"], n, n, attr=dev["attr"])
    r = ctx.hash_join(dev["fk"], dev["meas"], fk.size, desc, method="naive")
    assert r["n_match"] == fk.size - 8           # the 8 rows with fk 0 miss
    ref = O.single_join(O.DimSpec(keys, attr, n), fk, meas, method=O.NAIVE)
    assert r["n_match"] == ref["n_match"] and r["sum_product"] == ref["sum_product"]


# ------------------------------------------------------------------- Q3.3 plans --
@pytest.fixture(scope="module")
def q33_small():
    n = (1 << 20) + 37                      # ragged: not a multiple of 4 or 256
    host_dims = W.q33_host_dims(SEED)
    host_fact = W.q33_host_fact(SEED, n)
    return n, host_dims, host_fact, to_device({**host_dims, **host_fact})


@pytest.mark.parametrize("plan", W.q33_plans(), ids=[p[0] for p in W.q33_plans()])
def test_q33_plan_parity(ctx, q33_small, plan):
    name, method, order, bloom_order = plan
    _flush(f"q33 plan {name}")
    n, host_dims, host_fact, dev = q33_small
    res = ctx.star_join(W.q33_star_desc(dev, n, method, order, bloom_order))
    groups, lines, rows, oob = oracle_q33(host_dims, host_fact, order, method, bloom_order)
    assert gpu_groups(res) == groups
    assert res["n_groups"] == len(groups) > 0
    for (g, p, s, c, s32) in res["groups"]:
        assert s32 == ((s + 2**31) % 2**32) - 2**31     # reference int32 atomicAdd value
        assert g == p[0] * 1750 + p[1] * 7 + p[2] - 1992
    st = res["stats"]
    if method != "lip-adaptive":
        assert st["rows_in"][:4] == rows[:4]
        assert st["lines_in"][:4] == lines
        if method == "lip":
            assert st["rows_bloom_pass"] == rows[4]
    else:
        assert st["rows_in"][0] == n
        assert 0 < st["sample_rows"] <= n
        assert sorted(st["order_used"][:3]) == [0, 1, 2]


@pytest.mark.parametrize("order", [0, 2, 4])
@pytest.mark.parametrize("method", ["naive", "lip", "lip-adaptive"])
def test_q33_wide_predicates(ctx, q33_small, order, method):
    """Wider predicates: dense survivor segments, > 64 groups per workgroup
    (LDS table overflow to global atomics), tens of thousands of groups."""
    _flush(f"q33 wide {order} {method}")
    n, host_dims, host_fact, dev = q33_small
    city = ((L.COMB_SET, L.CMP_LT, 125),)
    year = ((L.COMB_SET, L.CMP_LTE, 1998),)
    dims = W.q33_dim_descs(dev, city_pred=city, year_pred=year)
    res = ctx.star_join(W.q33_star_desc(dev, n, method, W.Q33_ORDERS[order], dims=dims), groups_cap=1 << 17)
    groups, lines, rows, _ = oracle_q33(host_dims, host_fact, W.Q33_ORDERS[order], method,
                                        city_pred=city, year_pred=year)
    assert res["n_groups"] == len(groups) > 20000
    assert gpu_groups(res) == groups
    if method != "lip-adaptive":
        assert res["stats"]["lines_in"][:4] == lines


@pytest.mark.parametrize("n", [0, 1, 3, 4, 255, 256, 257, 1000, 4099])
def test_q33_tiny_and_ragged(ctx, n):
    host_dims = W.q33_host_dims(SEED)
    host_fact = W.q33_host_fact(SEED, max(n, 1))
    # make every row survive so small inputs still produce groups
    host_fact["lo_suppkey"][:] = 1 + np.arange(max(n, 1)) % 7
    host_fact["lo_custkey"][:] = 1 + np.arange(max(n, 1)) % 11
    city = ((L.COMB_SET, L.CMP_GTE, 0),)
    dev = to_device({**host_dims, **host_fact})
    dims = W.q33_dim_descs(dev, city_pred=city)
    for method in ("naive", "lip", "lip-adaptive"):
        res = ctx.star_join(W.q33_star_desc(dev, n, method, W.Q33_ORDERS[3], dims=dims))
        hf = {k: v[:n] for k, v in host_fact.items()}
        groups, _, _, _ = oracle_q33(host_dims, hf, W.Q33_ORDERS[3], method, city_pred=city)
        assert gpu_groups(res) == groups, (n, method)


def test_q33_resident_dims_match_inline_build(ctx, q33_small):
    n, host_dims, host_fact, dev = q33_small
    descs = W.q33_dim_descs(dev)
    dims = [ctx.build_dim(d) for d in descs]
    try:
        a = ctx.star_join(W.q33_star_desc(dev, n, "lip", W.Q33_ORDERS[0], resident=dims))
        b = ctx.star_join(W.q33_star_desc(dev, n, "lip", W.Q33_ORDERS[0]))
        assert a["groups"] == b["groups"] and a["stats"] == b["stats"]
        # a second query on the same resident tables gives the same answer
        c = ctx.star_join(W.q33_star_desc(dev, n, "naive", W.Q33_ORDERS[5], resident=dims))
        assert c["groups"] == a["groups"]
    finally:
        for d in dims:
            d.release()


# ---------------------------------------------------------------- error paths --
def test_out_of_domain_fact_keys_report_edomain(ctx, q33_small):
    n, host_dims, host_fact, dev = q33_small
    hf = {k: v.copy() for k, v in host_fact.items()}
    hf["lo_orderdate"][::1000] = 19900101        # below keys_min 19920101: reference UB, here a miss
    d2 = to_device(hf)
    d2.update({k: dev[k] for k in host_dims})
    for method in ("naive", "lip"):
        res = ctx.star_join(W.q33_star_desc(d2, n, method, W.Q33_ORDERS[2]), allow_domain=True)
        assert res["status"] == 6 and res["stats"]["n_oob"] > 0
        groups, _, _, _ = oracle_q33(host_dims, hf, W.Q33_ORDERS[2], method)
        assert gpu_groups(res) == groups
    with pytest.raises(L.LJError) as e:
        ctx.star_join(W.q33_star_desc(d2, n, "lip", W.Q33_ORDERS[2]))
    assert e.value.status == 6


def test_invalid_descriptors(ctx, q33_small):
    n, _, _, dev = q33_small
    bad = W.q33_star_desc(dev, n, "lip", (0, 0, 1))
    with pytest.raises(L.LJError) as e:
        ctx.star_join(bad)
    assert e.value.status == 1
    q = W.q33_star_desc(dev, n, "lip", (0, 1, 2))
    q.fact_fk[0] = q.fact_fk[0] + 4                # not 16-byte aligned
    with pytest.raises(L.LJError) as e:
        ctx.star_join(q)
    assert e.value.status == 1 and "aligned" in str(e.value)
    d = L.dim_desc(dev["s_suppkey"], 40000, 40000, attr=dev["s_city"], bloom=("mult", 1000, 2))
    with pytest.raises(L.LJError):
        ctx.build_dim(d)


def test_hbm_budget_is_enforced(q33_small):
    n, _, _, dev = q33_small
    with L.Context(0, hbm_budget=1 << 20) as small:
        with pytest.raises(L.LJError) as e:
            small.star_join(W.q33_star_desc(dev, n, "lip", (0, 1, 2)))
        assert e.value.status == 3
