"""Hash-table kinds on the GPU, through the C-ABI, against the oracle.

* PHT_1 (crystal/join.cuh:29-106,191-258): the key-only table, word for word,
  and joins (naive, LIP, adaptive; inline and resident) whose rows, sums,
  groups and per-stage statistics equal the oracle's.
* Textbook open addressing (key compare): the table holds exactly the
  selected keys (as a map: its layout depends on the order of the inserts
  under collisions), and joins equal the oracle's under heavy collisions, a
  full table and out-of-domain keys.
* The compat/textbook divergence of the survey's Appendix A1-A2 on the GPU:
  key 0 with value 0 and fact keys beyond the key range.
"""
import numpy as np
import pytest
import torch

import lipjoin as L
import oracle as O

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _groups(res):
    return [(g, p, s, c) for (g, p, s, c, _) in res["groups"]]


def _data(seed, nd, nf, kmin=0, fk_lo=None, fk_hi=None):
    rng = np.random.default_rng(seed)
    keys = (rng.permutation(nd) + kmin).astype(np.int32)
    attr = rng.integers(0, 40, nd).astype(np.int32)
    fk = rng.integers(kmin if fk_lo is None else fk_lo, kmin + nd if fk_hi is None else fk_hi, nf).astype(np.int32)
    meas = rng.integers(1, 100001, nf).astype(np.int32)
    return keys, attr, fk, meas


@pytest.mark.parametrize("kind", [L.HT_PHT1, L.HT_TEXTBOOK])
def test_table_words(ctx, kind):
    _flush(f"table words kind={kind}")
    nd = 300000
    keys, attr, _, _ = _data(21, nd, 1, kmin=1)
    keys[::1000] = 0                                   # key 0 (the PHT_1 sentinel) and repeats
    pred = ((L.COMB_SET, L.CMP_LT, 25),)
    cap = 1 << 19 if kind == L.HT_TEXTBOOK else nd + 1
    dk, da = _dev(keys), _dev(attr)
    dim = ctx.build_dim(L.dim_desc(dk, nd, cap, attr=da, predicate=pred, ht_kind=kind))
    ht, _ = dim.export()
    dim.release()
    flags = O.predicate(attr, [(O.COMB_SET, O.CMP_LT, 25)])
    if kind == L.HT_PHT1:
        ref, _ = O.pht1_build(keys, flags, cap)
        assert ht.size == cap and np.array_equal(ht, ref)
    else:
        assert ht.size == 2 * cap
        stored = ht[0::2] != 0
        got = dict(zip(((ht[0::2][stored].view(np.uint32) ^ np.uint32(0x80000000)).view(np.int32)).tolist(),
                       ht[1::2][stored].tolist()))
        ref_ht, _ = O.tb_build(keys, attr, flags, cap)
        rs = ref_ht[0::2] != 0
        ref = dict(zip(((ref_ht[0::2][rs].view(np.uint32) ^ np.uint32(0x80000000)).view(np.int32)).tolist(),
                       ref_ht[1::2][rs].tolist()))
        # a repeated key takes the last value: key 0 repeats, with values that may differ
        got.pop(0, None)
        ref.pop(0, None)
        assert got == ref


@pytest.mark.parametrize("kind,cap_of", [(L.HT_PHT1, lambda nd: nd), (L.HT_TEXTBOOK, lambda nd: 1 << 15),
                                         (L.HT_TEXTBOOK, lambda nd: 1 << 14)])
@pytest.mark.parametrize("method", ["naive", "lip", "lip-adaptive"])
def test_star_join_with_table_kind(ctx, kind, cap_of, method):
    """Two dimensions: the first of the given kind (group multiplier 0 for
    PHT_1, which has no payload), the second PHT_2."""
    nd, nf = 16384, (1 << 20) + 33
    cap = cap_of(nd)                                   # textbook: load factors 0.25 and 0.5
    _flush(f"star join kind={kind} cap={cap} {method}")
    k0, a0, f0, meas = _data(31, nd, nf, kmin=1)
    k1, a1, f1, _ = _data(32, 5000, nf, kmin=10)
    p0 = ((L.COMB_SET, L.CMP_LT, 20),)
    p1 = ((L.COMB_SET, L.CMP_GTE, 30),)
    gm = (0, 1, 0, 0) if kind == L.HT_PHT1 else (40, 1, 0, 0)
    bloom0 = ("mod", 1 << 14) if method != "naive" else None
    # the descriptors hold raw addresses: the tensors must outlive the call
    dk0, da0, dk1, da1, df0, df1, dm = (_dev(x) for x in (k0, a0, k1, a1, f0, f1, meas))
    d0 = L.dim_desc(dk0, nd, cap, attr=da0, predicate=p0, ht_kind=kind, bloom=bloom0)
    d1 = L.dim_desc(dk1, 5000, 5000, attr=da1, predicate=p1, keys_min=10,
                    bloom=("mod", 8192) if method != "naive" else None)
    q = L.star_desc([df0, df1], dm, nf, [d0, d1], (0, 1), method=method, group_mult=gm, n_groups=1600)
    res = ctx.star_join(q, groups_cap=2048)
    okind = {L.HT_PHT1: O.HT_PHT1, L.HT_TEXTBOOK: O.HT_TEXTBOOK}[kind]
    od = [O.DimSpec(k0, a0, cap, pred=[(O.COMB_SET, O.CMP_LT, 20)], ht_kind=okind,
                    bloom=O.bloom_spec(O.BLOOM_MOD, 1 << 14, 1) if method != "naive" else None),
          O.DimSpec(k1, a1, 5000, pred=[(O.COMB_SET, O.CMP_GTE, 30)], keys_min=10,
                    bloom=O.bloom_spec(O.BLOOM_MOD, 8192, 1, keys_min=10) if method != "naive" else None)]
    ref, lines, rows, _ = O.star_join(od, [f0, f1], meas, (0, 1), method=O.NAIVE if method == "naive" else O.LIP,
                                      group_mult=gm, n_groups=1600, threads=8)
    assert res["n_groups"] == len(ref) and _groups(res) == [(g, p, s, c) for g, p, s, c in ref]
    if method != "lip-adaptive":
        assert res["stats"]["lines_in"][:3] == lines and res["stats"]["rows_in"][:3] == rows[:3]


@pytest.mark.parametrize("kind", [L.HT_PHT1, L.HT_TEXTBOOK])
def test_single_join_table_kind_and_resident(ctx, kind):
    _flush(f"single join kind={kind}")
    nd, nf = 1 << 16, (1 << 21) + 5
    # fact keys over twice the key range: half of them are out of the dimension's domain
    keys, attr, fk, meas = _data(41, nd, nf, kmin=5, fk_lo=0, fk_hi=2 * nd)
    cap = nd if kind == L.HT_PHT1 else 1 << 17
    pred = ((L.COMB_SET, L.CMP_LT, 10),)
    dk, da, dfk, dm = _dev(keys), _dev(attr), _dev(fk), _dev(meas)
    okind = {L.HT_PHT1: O.HT_PHT1, L.HT_TEXTBOOK: O.HT_TEXTBOOK}[kind]
    spec = O.DimSpec(keys, attr, cap, pred=[(O.COMB_SET, O.CMP_LT, 10)], keys_min=5, ht_kind=okind)
    ref = O.single_join(spec, fk, meas, method=O.NAIVE, threads=8)
    desc = L.dim_desc(dk, nd, cap, attr=da, predicate=pred, keys_min=5, ht_kind=kind)
    for method in ("naive", "lip"):
        r = ctx.hash_join(dfk, dm, nf, desc, method=method, allow_domain=True)
        assert (r["n_match"], r["sum_measure"], r["sum_product"]) == (ref["n_match"], ref["sum_measure"],
                                                                      ref["sum_product"]), method
    dim = ctx.build_dim(desc)
    try:
        for method in ("naive", "lip"):
            q = L.star_desc([dfk], dm, nf, None, (0,), method=method, resident=[dim])
            r = ctx.star_join(q, allow_domain=True)
            assert (r["n_rows"], r["sum_total"]) == (ref["n_match"], ref["sum_measure"])
    finally:
        dim.release()


def test_compat_and_textbook_diverge_as_documented(ctx):
    """Key 0 with value 0 and fact keys beyond the key range (Appendix A1-A2)."""
    _flush("compat vs textbook divergence")
    nd = 1 << 12
    keys = np.arange(nd, dtype=np.int32)
    vals = ((keys * 7) % 13).astype(np.int32)          # key 0 has value 0
    nf = 1 << 18
    rng = np.random.default_rng(51)
    fk = rng.integers(0, 2 * nd, nf).astype(np.int32)  # half beyond the range: alias in PHT_2
    meas = rng.integers(1, 1000, nf).astype(np.int32)
    dk, dv, dfk, dm = _dev(keys), _dev(vals), _dev(fk), _dev(meas)
    out = {}
    for kind, cap, okind in ((L.HT_PHT2, nd, O.HT_PHT2), (L.HT_TEXTBOOK, 2 * nd, O.HT_TEXTBOOK)):
        r = ctx.hash_join(dfk, dm, nf, L.dim_desc(dk, nd, cap, attr=dv, ht_kind=kind), method="naive")
        ref = O.single_join(O.DimSpec(keys, vals, cap, ht_kind=okind), fk, meas, method=O.NAIVE)
        assert (r["n_match"], r["sum_measure"], r["sum_product"]) == \
            (ref["n_match"], ref["sum_measure"], ref["sum_product"])
        out[kind] = r
    in_dom = (fk < nd)
    # textbook: exactly the in-range keys match (key 0 included)
    assert out[L.HT_TEXTBOOK]["n_match"] == int(in_dom.sum())
    # compat: every key but the ones on key 0's all-zero slot (fk % nd == 0)
    assert out[L.HT_PHT2]["n_match"] == int((fk % nd != 0).sum())


@pytest.mark.parametrize("method", ["lip", "lip-adaptive"])
def test_filter_bits_past_the_staged_prefix(ctx, method):
    """A modulo filter is staged in LDS up to the bits the table's key domain
    can set (ht_len); a dimension key beyond the domain sets a bit past that
    prefix, and fact keys that reach it read the filter's global words.  The
    results must still equal the oracle's (the PHT aliases such keys)."""
    _flush(f"filter prefix {method}")
    nd, nf = 3000, (1 << 19) + 7
    rng = np.random.default_rng(61)
    keys = (rng.permutation(nd) + 1).astype(np.int32)
    # 40 keys moved beyond the domain by ht_len: the same PHT slot (no
    # collision: the insert order cannot matter), a filter bit past the
    # staged prefix (bits 3100..6000 of 8192; the prefix covers 3072)
    moved = np.nonzero(keys > 100)[0][:40]
    keys[moved] += nd
    attr = rng.integers(0, 10, nd).astype(np.int32)
    k1 = (rng.permutation(2000) + 1).astype(np.int32)
    a1 = rng.integers(0, 10, 2000).astype(np.int32)
    fk = rng.integers(1, 8192, nf).astype(np.int32)                # all of the filter's bits are reached
    f1 = rng.integers(1, 2001, nf).astype(np.int32)
    meas = rng.integers(1, 1000, nf).astype(np.int32)
    dk, da, dk1, da1, dfk, df1, dm = (_dev(x) for x in (keys, attr, k1, a1, fk, f1, meas))
    pred = ((L.COMB_SET, L.CMP_LT, 6),)
    d0 = L.dim_desc(dk, nd, nd, attr=da, predicate=pred, bloom=("mod", 8192))
    d1 = L.dim_desc(dk1, 2000, 2000, attr=da1, predicate=pred, bloom=("mod", 2048))
    q = L.star_desc([dfk, df1], dm, nf, [d0, d1], (0, 1), method=method, group_mult=(10, 1, 0, 0), n_groups=100)
    res = ctx.star_join(q, groups_cap=256)
    od = [O.DimSpec(keys, attr, nd, pred=[(O.COMB_SET, O.CMP_LT, 6)], bloom=O.bloom_spec(O.BLOOM_MOD, 8192, 1)),
          O.DimSpec(k1, a1, 2000, pred=[(O.COMB_SET, O.CMP_LT, 6)], bloom=O.bloom_spec(O.BLOOM_MOD, 2048, 1))]
    ref, lines, rows, _ = O.star_join(od, [fk, f1], meas, (0, 1), method=O.LIP, group_mult=(10, 1, 0, 0),
                                      n_groups=100, threads=8)
    assert _groups(res) == [(g, p, s, c) for g, p, s, c in ref]
    assert res["stats"]["rows_bloom_pass"] == rows[3]
