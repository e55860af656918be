"""The oracle's star join on the SSB flight 2/3 queries (predicate steps,
payload columns and group encodings of src/ssb-jo/q{21..34}) against an
independent numpy evaluation of each query's SQL meaning."""
import pytest

from lipjoin import ssb
from lipjoin import workloads as W
from ssb_helpers import SF, brute_force, canon, oracle_query

N_FACT = 1 << 18


@pytest.fixture(scope="module")
def data():
    seed = W.seed_of(6)
    return ssb.host_dims(seed, SF), ssb.host_fact(seed, N_FACT, SF)


def test_dimension_encodings(data):
    h, f = data
    c_len, s_len, p_len, d_len = ssb.table_sizes(SF)
    assert h["c_key"].size == c_len and h["s_key"].size == s_len and h["p_key"].size == p_len
    assert (h["s_region"] == h["s_nation"] // 5).all() and (h["c_city"] // 10 == h["c_nation"]).all()
    assert (h["p_brand1"] // 40 == h["p_category"]).all() and h["p_brand1"].max() < 1000
    assert set(h["d_year"].tolist()) == set(range(1992, 1999))
    assert h["d_yearmonthnum"].min() == 199201 and h["d_yearmonthnum"].max() == 199812
    assert (h["p_mfgr"] == h["p_category"] // 5).all()
    assert f["lo_partkey"].min() >= 1 and f["lo_partkey"].max() <= p_len


@pytest.mark.parametrize("name", sorted(ssb.QUERIES))
@pytest.mark.parametrize("method,order", [("naive", 0), ("lip", 0), ("naive", 3), ("lip", 4)])
def test_oracle_matches_sql_meaning(data, name, method, order):
    h, f = data
    q = ssb.QUERIES[name]
    orders = ssb.orders_of(q)
    if order not in orders:
        order = 0
    groups, _, rows, oob = oracle_query(q, h, f, orders[order], method)
    want = brute_force(name, h, f)
    assert oob == 0
    assert len(want) > 0, "the query selects no row at this size"
    assert canon(q, groups) == canon(q, want)


def test_q41_other_join_order(data):
    """The engine takes any order of the four dimensions; one more than the reference ships."""
    h, f = data
    q = ssb.QUERIES["q41"]
    for method in ("naive", "lip"):
        got = oracle_query(q, h, f, (3, 2, 0, 1), method)[0]
        assert canon(q, got) == canon(q, brute_force("q41", h, f))


def test_column_files_and_plan_names(tmp_path):
    """The CLI reads the reference's column files (ssb_utils.h naming) and its plan names."""
    h = ssb.host_dims(W.seed_of(6), SF)
    h["s_key"].astype("<i4").tofile(tmp_path / "SUPPLIER0")
    h["s_region"].astype("<i4").tofile(tmp_path / "SUPPLIER5")
    got = ssb.read_columns(str(tmp_path), ["s_key", "s_region"])
    assert (got["s_key"] == h["s_key"]).all() and (got["s_region"] == h["s_region"]).all()
    assert ssb.parse_plan(ssb.QUERIES["q33"], "3-lip") == ("lip", ssb.ORDERS[3])
    assert ssb.parse_plan(ssb.QUERIES["q42"], "0-naive") == ("naive", (0, 1, 2, 3))
    assert ssb.parse_plan(ssb.QUERIES["q21"], "lip-adaptive")[0] == "lip-adaptive"
    for bad in ("6-lip", "1-lip", "x", "0-bloom"):
        q = ssb.QUERIES["q41" if bad == "1-lip" else "q21"]
        with pytest.raises(ValueError):
            ssb.parse_plan(q, bad)
    assert set(ssb.PRINT_ORDER) == set(ssb.QUERIES)
