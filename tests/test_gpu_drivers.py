"""The host drivers that stand in for the reference's binaries, run as
processes on the GPU and checked against the oracle:

* bin/lj_join (bin/ops/join and bin/ops/bloom_join): the reference's own
  generator (srand(0), generator.h:37-136), predicate dim_val <= num_select
  (bloom_join.cu:61), plain and Bloom joins; its JSON line's matches and
  sums must equal the oracle's on the same inputs.
* bin/lj_q33 (bin/ssb-jo/q33/<plan>): the seeded Q3.3 columns; its "Result:"
  rows (year, c_city, s_city, sum) must equal the oracle's groups.
"""
import json
import os
import subprocess

import pytest

import lipjoin as L
import oracle as O
from lipjoin import workloads as W
from q33_helpers import oracle_q33

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "crystal-opt_amd", "bin")


def _run(args):
    print("run", " ".join(args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout


@pytest.mark.parametrize("bits", [None, 8192, 3000])
def test_lj_join_reference_generator(bits):
    n, d, s = 300_007, 4096, 1000
    args = [os.path.join(BIN, "lj_join"), f"--n={n}", f"--d={d}", f"--s={s}", "--t=2"]
    if bits:
        args.append(f"--b={bits}")
    lines = [json.loads(x) for x in _run(args).splitlines() if x.startswith("{")]
    assert len(lines) == 2
    g = L.gen_reference_host(d, n)
    bloom = O.bloom_spec(O.BLOOM_MOD, bits, 1) if bits else None
    spec = O.DimSpec(g["dim_keys"], g["dim_vals"], d, pred=[(O.COMB_SET, O.CMP_LTE, s)], bloom=bloom)
    ref = O.single_join(spec, g["fact_keys"], g["fact_vals"], method=O.LIP if bits else O.NAIVE, threads=8)
    for x in lines:
        assert (x["n_match"], x["sum"], x["sum_measure"]) == (ref["n_match"], ref["sum_product"],
                                                            ref["sum_measure"])
        assert x["num_bloom_filter_bits"] == (bits or 0) and x["time_probe"] > 0


@pytest.mark.parametrize("plan", ["0-lip", "3-naive", "3-lip-best-bloom"])
def test_lj_q33_result_rows(plan):
    n = 2_000_003
    out = _run([os.path.join(BIN, "lj_q33"), f"--plan={plan}", "--t=1", f"--n={n}"])
    lines = out.splitlines()
    i = lines.index("Result:")
    rows = []
    for x in lines[i + 1:]:
        if x.startswith("Res Count"):
            break
        rows.append(tuple(int(v) for v in x.split()))
    seed = W.seed_of(3)
    name, method, order, bo = [p for p in W.q33_plans() if p[0] == plan][0]
    groups, _, _, _ = oracle_q33(W.q33_host_dims(seed), W.q33_host_fact(seed, n), order, method, bo)
    # the reference prints year, c_city, s_city, sum (q33/0-lip.cu:288-295)
    assert rows == [(p[2], p[1], p[0], s) for (g, p, s, c) in groups]
    assert any(x.startswith('{"query":33,"time_query":') for x in lines)
