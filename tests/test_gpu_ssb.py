"""GPU parity of the SSB flight 2/3 queries (src/ssb-jo/q{21,22,23,31,32,33,34}):
every join order x {naive, LIP, adaptive LIP} through lj_star_join against the
oracle on the same seeded SF-1 tables.  Integer work: group keys, payloads,
int64 sums, counts, per-stage line statistics and the reference's int32
atomicAdd value are compared exactly."""
import pytest
import torch

from lipjoin import ssb
from lipjoin import workloads as W
from q33_helpers import gpu_groups, to_device
from ssb_helpers import SF, brute_force, canon, oracle_query

pytestmark = pytest.mark.gpu

N_FACT = (1 << 20) + 77          # a ragged last tile


@pytest.fixture(scope="module")
def sctx():
    """A context of its own: its group accumulators are sized by these queries'
    group spaces (<= 8192 groups: eight copies, one per XCD), not by the
    437,500-group Q3.3 space the shared context has seen."""
    import lipjoin as L
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = L.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ssb_data():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    seed = W.seed_of(6)
    h, f = ssb.host_dims(seed, SF), ssb.host_fact(seed, N_FACT, SF)
    return h, f, to_device({**h, **f})


def _wrap32(x):
    return (x + 2**31) % 2**32 - 2**31


@pytest.mark.parametrize("name", sorted(ssb.QUERIES))
def test_ssb_query_all_plans(sctx, ssb_data, name):
    ctx = sctx
    h, f, dev = ssb_data
    q = ssb.QUERIES[name]
    dims = ssb.dim_descs(q, dev, SF)
    want = canon(q, brute_force(name, h, f))
    assert want
    orders = dict(ssb.orders_of(q))
    if len(q.dims) == 4:
        orders["extra"] = (3, 2, 0, 1)
    for oi, order in orders.items():
        for method in ("naive", "lip", "lip-adaptive"):
            print(f"ssb {name} order {oi} {method}", flush=True)
            res = ssb.run(ctx, q, dev, N_FACT, method, order, SF, dims=dims)
            assert res["n_groups"] == len(want), (name, oi, method)
            assert canon(q, gpu_groups(res)) == want, (name, oi, method)
            assert all(g[4] == _wrap32(g[2]) for g in res["groups"]), "int32 atomicAdd value"
            if method != "lip-adaptive":
                _, lines, _, oob = oracle_query(q, h, f, order, method)
                assert oob == 0 and res["stats"]["lines_in"][:len(lines)] == lines, (name, oi, method)


@pytest.mark.parametrize("name", ["q21", "q31", "q34", "q41"])
def test_ssb_resident_dims(sctx, ssb_data, name):
    ctx = sctx
    """Tables built once (lj_dim_build) serve every plan of the query."""
    h, f, dev = ssb_data
    q = ssb.QUERIES[name]
    built = [ctx.build_dim(d) for d in ssb.dim_descs(q, dev, SF)]
    try:
        want = canon(q, brute_force(name, h, f))
        order = ssb.ORDERS[2] if len(q.dims) == 3 else (2, 0, 3, 1)
        for method in ("naive", "lip"):
            res = ssb.run(ctx, q, dev, N_FACT, method, order, SF, resident=built)
            assert canon(q, gpu_groups(res)) == want, (name, method)
    finally:
        for d in built:
            d.release()


def test_group_copies_follow_the_group_space(ssb_data):
    """A fresh context keeps eight accumulator copies for small group spaces and
    one for a large one (re-sized when a larger space arrives); every result is
    exact before and after the re-size."""
    import lipjoin as L
    h, f, dev = ssb_data
    q = ssb.QUERIES["q31"]
    want = canon(q, brute_force("q31", h, f))
    with L.Context(0) as c:
        for big in (False, True):
            if big:
                # a 437,500-group query (Q3.3's group space) over the same rows
                q33 = ssb.QUERIES["q33"]
                r = ssb.run(c, q33, dev, N_FACT, "lip", ssb.ORDERS[0], SF)
                assert canon(q33, gpu_groups(r)) == canon(q33, brute_force("q33", h, f))
            for method in ("naive", "lip-adaptive"):
                res = ssb.run(c, q, dev, N_FACT, method, ssb.ORDERS[0], SF)
                assert canon(q, gpu_groups(res)) == want, (big, method)


def test_measure_sub_errors(sctx, ssb_data):
    """fact_measure_sub needs two or more dimensions and 16-byte alignment."""
    import lipjoin as L
    h, f, dev = ssb_data
    q = ssb.QUERIES["q41"]
    desc = ssb.star(q, dev, N_FACT, "lip", (0, 1, 2, 3), SF)
    desc.fact_measure_sub = dev["lo_supplycost"].data_ptr() + 4
    with pytest.raises(L.LJError) as e:
        sctx.star_join(desc)
    assert e.value.status == 1
    one = L.star_desc([dev["lo_suppkey"]], dev["lo_revenue"], N_FACT, ssb.dim_descs(q, dev, SF)[:1], (0,),
                      method="lip", group_mult=(0,), n_groups=1)
    one.fact_measure_sub = dev["lo_supplycost"].data_ptr()
    with pytest.raises(L.LJError) as e:
        sctx.star_join(one)
    assert e.value.status == 5


@pytest.mark.parametrize("name,plan", [("q31", "2-naive"), ("q22", "5-lip"), ("q43", "0-lip")])
def test_cli_reads_reference_column_files(ssb_data, tmp_path, capsys, name, plan):
    """python -m lipjoin.ssb over files in the reference's column format prints
    the reference's result rows (payloads in its order, then the sum)."""
    h, f, _ = ssb_data
    q = ssb.QUERIES[name]
    for k, fname in ssb.COLUMN_FILES.items():
        src = h.get(k, f.get(k))
        src.astype("<i4").tofile(tmp_path / fname)
    assert ssb.main([f"--query={name}", f"--plan={plan}", "--t=1", f"--data_dir={tmp_path}"]) == 0
    out = capsys.readouterr().out.splitlines()
    rows = out[out.index("Result:") + 1: next(i for i, l in enumerate(out) if l.startswith("Res Count"))]
    want = [" ".join(str(g[1][i]) for i in ssb.PRINT_ORDER[name]) + f" {g[2]}"
            for g in canon(q, brute_force(name, h, f))]
    assert rows == want
    assert out[-1].startswith('{"query":%s,"time_query":' % name[1:])


def test_ssb_batch_matches_brute_force(sctx, ssb_data):
    """One lj_batch over SSB queries of three and four dimensions (Q4.x with the
    second measure column, adaptive plans with the stage-0 mask): every
    query's groups equal its brute force."""
    h, f, dev = ssb_data
    plans = [("q21", "lip", 0), ("q31", "lip-adaptive", 0), ("q41", "lip-adaptive", 0), ("q42", "naive", 0),
             ("q43", "lip", 0), ("q43", "lip-adaptive", 0), ("q33", "lip", 3)]
    descs = []
    for name, method, oi in plans:
        q = ssb.QUERIES[name]
        d = ssb.star(q, dev, N_FACT, method, ssb.orders_of(q)[oi], SF)
        if q.measure_sub:
            d.fact_measure_sub = dev[q.measure_sub].data_ptr()
        descs.append(d)
    b = sctx.batch(descs, groups_cap=1 << 12)
    try:
        for rep in range(2):
            res = b.run()
            for (name, method, oi), r in zip(plans, res):
                q = ssb.QUERIES[name]
                assert canon(q, gpu_groups(r)) == canon(q, brute_force(name, h, f)), (rep, name, method)
    finally:
        b.close()
