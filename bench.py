#!/usr/bin/env python3
"""Benchmark: SSB Q3.3-shaped LIP star join on one MI355X.

Workload: 120,000,000-row fact table (lo_custkey, lo_suppkey, lo_orderdate,
lo_partkey, lo_revenue), customer 600,000 / supplier 40,000 / date 2,556 rows,
s_city, c_city in {231, 235}, 1992 <= d_year <= 1997, GROUP BY (c_city, s_city,
d_year) with int64 SUM(lo_revenue).  Synthetic, seeded, generated on the device.

A step is one whole query, the reference's runQuery window: table fills,
the three dimension builds, the fused probe and the group finalize (the
results are copied back and checked each step).  Timed steps use the
adaptive-LIP plan; every one of the reference's 13 plans (6 join orders x
{naive, LIP} + best-bloom) is also timed and reported under "plans".

Roofline: the fused probe kernel is the dominant kernel.  Its algorithmic
bytes per launch are 128 B x the 128-byte lines of fact columns that hold at
least one row still alive when the column is read (exact counts reported by
the kernel, equal to the oracle's in the parity tests), using the minimum
over the 6 static LIP orders; achieved = bytes / average probe-kernel time
(HIP events on the kernel's stream); peak = 8 TB/s HBM3E.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "crystal-opt_amd"))

METRIC = "fact rows/s + %HBM peak, LIP vs naive star join; batched p50/p99 latency"
HBM_PEAK = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per probe launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE of the probe kernel, each scaled by the same
    counter's reading for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "probe_q33_120M_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def bench_batch(ctx, n_inst, reps):
    """Config 4: n_inst single joins (fact rows 2^20..2^27 log-uniform, 2^22-key
    dimension each, selectivity 0.5-20 %, Bloom filter from the autotuner) as one
    batch (graph replay and eager) against n_inst separate lj_star_join calls."""
    import torch
    from lipjoin import workloads as W
    ctx.tune_calibrate()
    specs = W.batch_specs(W.seed_of(4), n_inst=n_inst)
    cols, descs, choices = [], [], []
    for sp in specs:
        c = W.batch_instance_device(ctx, sp)
        bloom, ch = W.tuned_bloom(ctx, sp, semi_join=True)   # one group, no product
        cols.append(c)
        choices.append(ch)
        descs.append(W.batch_desc(c, sp, method="lip", bloom=bloom))
    torch.cuda.synchronize()
    rows = sum(sp["n_fact"] for sp in specs)
    sep = [ctx.star_join(d) for d in descs]          # warm-up and reference results
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for d in descs:
            ctx.star_join(d)
        t.append(time.perf_counter() - t0)
    sep_s = statistics.median(t)
    out = {"instances": n_inst, "fact_rows_total": rows, "separate_ms": sep_s * 1e3}
    # algorithmic bytes of one batch: per instance the FK column (4 B/row), the
    # 128-byte measure lines holding a survivor, and the dimension build (key
    # and attr read, 8 B/row; 8 B per inserted slot; the filter's words)
    dim_bytes = sum(8 * sp["n_dim"] for sp in specs)
    dim_bytes += sum(8 * round(sp["n_dim"] * sp["threshold"] / 10000.0) + ch["bloom_bits"] // 8
                     for sp, ch in zip(specs, choices))
    fact_bytes = sum(4 * sp["n_fact"] + 128 * r["stats"]["lines_in"][1] for sp, r in zip(specs, sep))
    out["bytes_alg"] = fact_bytes + dim_bytes
    for mode, graph, per_kind in (("eager", False, False), ("graph", True, False), ("eager_per_kind", False, True)):
        b = ctx.batch(descs, use_graph=graph, launch_per_kind=per_kind)
        res = b.run()
        exact = all((r["sum_total"], r["n_rows"]) == (s["sum_total"], s["n_rows"]) for r, s in zip(res, sep))
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            res = b.run()
            t.append(time.perf_counter() - t0)
        exact = exact and all((r["sum_total"], r["n_rows"]) == (s["sum_total"], s["n_rows"])
                              for r, s in zip(res, sep))
        b.close()
        ms = statistics.median(t) * 1e3
        out[f"batch_{mode}_ms"] = ms
        out[f"batch_{mode}_rows_per_s"] = rows / (ms * 1e-3)
        out[f"batch_{mode}_hbm_frac"] = out["bytes_alg"] / (ms * 1e-3) / HBM_PEAK
        out[f"speedup_{mode}_vs_separate"] = sep_s * 1e3 / ms
        out[f"bit_exact_{mode}"] = exact
    kinds = {}
    for ch in choices:
        k = f"{'mod' if ch['bloom_hash'] == 1 else 'mult'}-k{ch['bloom_k']}-{'blk' if ch['bloom_layout'] else 'std'}" \
            f"-2^{int(ch['bloom_bits']).bit_length() - 1}{'-lds' if ch['in_lds'] else ''}"
        kinds[k] = kinds.get(k, 0) + 1
    out["bloom_choices"] = kinds
    out["tune_model"] = ctx.tune_model()
    out["separate_rows_per_s"] = rows / sep_s
    log(f"batch: {n_inst} joins, {rows / 1e9:.2f} G rows: separate {sep_s * 1e3:.2f} ms, "
        f"eager {out['batch_eager_ms']:.2f} ms, graph {out['batch_graph_ms']:.2f} ms, "
        f"a launch per kind {out['batch_eager_per_kind_ms']:.2f} ms; filters {kinds}; "
        f"tuner model {out['tune_model']}")
    del cols
    return out


def bench_single(ctx, reps=5):
    """Configs 1 and 2: the plain FK-PK join (2^20-row dimension, 2^24 fact
    rows, attr < 10 of U[0,99], no filter) and the Bloom join (2^22 / 2^28,
    selectivity 1-50 %, MULT filters 2^16..2^26 bits x k 1..3 and the
    autotuner's choice, LIP against naive).  Probe times from HIP events;
    algorithmic bytes = 4 B x fact rows + 128 B x measure lines touched."""
    import torch
    from lipjoin import workloads as W

    def timed(dfk, dm, nf, desc, method, sum_only=False):
        r = None
        ts = []
        for i in range(reps + 1):
            r = ctx.hash_join(dfk, dm, nf, desc, method=method, sum_only=sum_only)
            if i:
                ts.append(r["time_probe"])
        st = r["stats"]
        b = 4 * nf + 128 * sum(st["lines_in"][1:2])
        ms = statistics.median(ts)
        return {"probe_ms": ms, "rows_per_s": nf / (ms * 1e-3), "bytes_alg": b,
                "hbm_frac": b / (ms * 1e-3) / HBM_PEAK, "n_match": r["n_match"], "sum": r["sum_measure"]}

    out = {}
    seed, nd, nf = W.seed_of(1), 1 << 20, 1 << 24
    dk, da, dfk, dm = W.single_join_device(ctx, seed, nd, nf, 99)
    desc = L_dim(dk, nd, da, W.single_join_pred(10), None)
    out["config1"] = timed(dfk, dm, nf, desc, "naive")
    del dk, da, dfk, dm
    seed, nd, nf = W.seed_of(2), 1 << 22, 1 << 28
    dk, da, dfk, dm = W.single_join_device(ctx, seed, nd, nf, 9999)
    torch.cuda.synchronize()
    rows = []
    for sel in (1, 5, 10, 25, 50):
        thr = 100 * sel
        pred = W.single_join_pred(thr)
        e = {"selectivity_pct": sel, "naive": timed(dfk, dm, nf, L_dim(dk, nd, da, pred, None), "naive")}
        best = None
        for lg in (16, 18, 20, 22, 24, 26):
            for k in (1, 2, 3):
                r = timed(dfk, dm, nf, L_dim(dk, nd, da, pred, ("mult", 1 << lg, k, 0)), "lip")
                if best is None or r["probe_ms"] < best[1]["probe_ms"]:
                    best = (f"mult k={k} 2^{lg}", r)
        e["lip_best"] = dict(best[1], bloom=best[0])
        sp = {"n_dim": nd, "threshold": thr}
        bl, ch = W.tuned_bloom(ctx, sp)
        e["lip_tuned"] = dict(timed(dfk, dm, nf, L_dim(dk, nd, da, pred, bl), "lip"), bloom=str(bl))
        # the tuned filter's words: popcount and a 64-bit hash (bit-exactness
        # against the oracle is tested word for word)
        dim = ctx.build_dim(L_dim(dk, nd, da, pred, bl))
        _, words = dim.export()
        e["lip_tuned"]["bloom_popcount"] = dim.info()["bloom_bits_set"]
        e["lip_tuned"]["bloom_words_xxh64"] = words_hash(words)
        dim.release()
        # SUM(measure) only (the configuration's output; no product): the
        # filter tuned for a semi-join, whose exact modulo filter decides the
        # table probe
        bs, _ = W.tuned_bloom(ctx, sp, semi_join=True)
        e["lip_semi"] = dict(timed(dfk, dm, nf, L_dim(dk, nd, da, pred, bs), "lip", sum_only=True), bloom=str(bs))
        assert e["lip_tuned"]["sum"] == e["naive"]["sum"] == e["lip_best"]["sum"] == e["lip_semi"]["sum"]
        assert e["lip_tuned"]["n_match"] == e["lip_semi"]["n_match"]
        rows.append(e)
        log(f"config2 s={sel}%: naive {e['naive']['probe_ms']:.3f} ms, LIP best {e['lip_best']['probe_ms']:.3f} "
            f"({best[0]}), tuned {e['lip_tuned']['probe_ms']:.3f} ({bl}), sum-only {e['lip_semi']['probe_ms']:.3f} "
            f"({bs})")
    out["config2"] = rows
    return out


def words_hash(words):
    """XXH64 of a filter's words (little-endian bytes), hex."""
    try:
        import xxhash
        return xxhash.xxh64(words.tobytes()).hexdigest()
    except ImportError:
        import hashlib
        return "blake2b-" + hashlib.blake2b(words.tobytes(), digest_size=8).hexdigest()


def bench_bf_size(ctx, reps=3):
    """The reference's one logged measurement of this path (eval/plot_bf_size.py:
    110-115): bin/ops/bloom_join --n=2^30 --d=524288 --s=4096 --b=4096, i.e. a
    2^30-row fact table over a 524,288-row dimension, dim_val <= 4096 selected,
    a 4,096-bit modulo filter; time_probe 20.1574 / 20.0898 / 20.098 ms on
    unstated hardware.  Inputs from the reference's own generator (srand(0),
    generator.h:37-136) on the host; the same join through lj_hash_join, its
    counts and sums checked against the oracle."""
    import torch
    import lipjoin as L
    n, d, sel, bits = 1 << 30, 524288, 4096, 4096
    log(f"bf_size: generating {n} fact rows with the reference generator (host rand())")
    g = L.gen_reference_host(d, n, want=("dim_keys", "dim_vals", "fact_keys"))
    dk, dv = torch.from_numpy(g["dim_keys"]).cuda(), torch.from_numpy(g["dim_vals"]).cuda()
    fk = torch.from_numpy(g["fact_keys"]).cuda()
    fv = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.gen_device(fv, L.GEN_SEQUENCE, 0, 0, 0, 0)          # dummy_initialize: 0..n-1
    torch.cuda.synchronize()
    desc = L.dim_desc(dk, d, d, attr=dv, predicate=((L.COMB_SET, L.CMP_LTE, sel),), bloom=("mod", bits))
    runs = [ctx.hash_join(fk, fv, n, desc, method="lip") for _ in range(reps + 1)][1:]
    r = runs[-1]
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    spec = O.DimSpec(g["dim_keys"], g["dim_vals"], d, pred=[(O.COMB_SET, O.CMP_LTE, sel)],
                     bloom=O.bloom_spec(O.BLOOM_MOD, bits, 1))
    ref = O.single_join(spec, g["fact_keys"], np.arange(n, dtype=np.int32), method=O.LIP,
                        threads=host_cpu_share()[0])
    probe = statistics.median(x["time_probe"] for x in runs)
    out = {"workload": "bloom_join --n=2^30 --d=524288 --s=4096 --b=4096 (reference generator inputs)",
           "time_probe_ms": probe, "time_build_ms": statistics.median(x["time_build"] for x in runs),
           "time_extra_ms": statistics.median(x["time_extra"] for x in runs),
           "time_join_total_ms": statistics.median(x["time_total"] for x in runs),
           "fact_rows_per_s": n / (probe * 1e-3), "n_match": r["n_match"], "n_bloom_pass": r["n_bloom_pass"],
           "reference_logged_time_probe_ms": [20.1574, 20.0898, 20.098],
           "probe_speedup_vs_reference_log": 20.098 / probe,
           "matches_oracle": (r["n_bloom_pass"], r["n_match"], r["sum_measure"], r["sum_product"]) ==
                             (ref["n_bloom_pass"], ref["n_match"], ref["sum_measure"], ref["sum_product"])}
    log(f"bf_size: probe {probe:.3f} ms (reference log 20.1 ms), matches oracle {out['matches_oracle']}")
    del g, dk, dv, fk, fv
    torch.cuda.empty_cache()
    return out


def L_dim(keys, n, attr, pred, bloom):
    import lipjoin as L
    return L.dim_desc(keys, n, n, attr=attr, predicate=pred, bloom=bloom)


def _pct(xs, p):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(p / 100.0 * (len(xs) - 1))))] if xs else None


def bench_serve(ctx, n_req, fractions=(0.25, 0.5, 0.75, 0.9, 1.1, 1.5, 2.0)):
    """Config 5: Poisson arrivals of single-join requests (fact slices of
    2^18..2^26 rows from a resident 2^26-row pool, one of 8 predicate templates
    over a resident 2^22-key dimension) at rates that are fractions of the
    measured one-at-a-time capacity; p50/p99 latency (arrival to result ready)
    and rows/s for the serving loop and for one-query-at-a-time."""
    import numpy as np
    import torch
    from lipjoin import workloads as W
    seed = W.seed_of(5)
    n_pool = 1 << 26
    cols, tmpls = W.serve_templates(ctx, seed)
    pool = W.serve_pool(ctx, seed, n_pool)
    torch.cuda.synchronize()
    reqs = W.serve_requests(seed, n_req, len(tmpls), n_pool)
    descs = [W.serve_desc(pool, tmpls[t], o, n) for (t, o, n) in reqs]
    rows = [n for (_, _, n) in reqs]
    # every request's result alone (also the warm-up): (status, rows, sum)
    ref = []
    for d in descs:
        st = ctx.star_join_raw(d)
        ref.append((st, ctx._raw_res.n_rows, ctx._raw_res.sum_total))
    # a sample of them against the oracle's single join on the same slice
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    keys, attr = cols["keys"].cpu().numpy(), cols["attr"].cpu().numpy()
    fk_h, meas_h = pool["fk"][:n_pool].cpu().numpy(), pool["meas"][:n_pool].cpu().numpy()
    oracle_ok, oracle_n = True, 0
    for i in range(0, n_req, max(1, n_req // 16)):
        t_i, o, nr = reqs[i]
        b = tmpls[t_i]["bloom"]
        bs = (O.bloom_spec(O.BLOOM_MOD, b[1], 1) if b[0] == "mod" else
              O.bloom_spec(O.BLOOM_MULT, b[1], b[2], layout=b[3]))
        spec = O.DimSpec(keys, attr, keys.size, pred=list(W.single_join_pred(tmpls[t_i]["threshold"])), bloom=bs)
        r_o = O.single_join(spec, fk_h[o:o + nr], meas_h[o:o + nr], method=O.LIP, threads=os.cpu_count() or 1)
        oracle_ok = oracle_ok and ref[i] == (0, r_o["n_match"], r_o["sum_measure"])
        oracle_n += 1
    del fk_h, meas_h
    t = []
    for d in descs[:min(n_req, 500)]:
        t0 = time.perf_counter()
        ctx.star_join(d)
        t.append(time.perf_counter() - t0)
    service = statistics.fmean(t)
    capacity = 1.0 / service
    out = {"requests_per_rate": n_req, "serial_service_ms": service * 1e3, "serial_capacity_req_s": capacity,
           "hbm_budget_bytes": 200 << 30, "oracle_sample": oracle_n, "oracle_sample_match": oracle_ok,
           "rates": []}
    rng = np.random.default_rng(seed + 7)
    for f in fractions:
        lam = f * capacity
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
        # serving loop
        srv = ctx.serve(hbm_budget=200 << 30, max_batch=64, max_inflight=4, prewarm_groups=1)
        lag = []
        t0 = time.perf_counter()
        ids = []
        for i, d in enumerate(descs):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            lag.append(time.perf_counter() - t0 - arrivals[i])
            ids.append(srv.submit(d))
        res = [srv.poll(i, timeout_ms=120000) for i in ids]
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.stop()
        lat = [r["time_query_ms"] + 1e3 * lg for r, lg in zip(res, lag)]
        # every request: served result == the same request alone
        exact = all(r is not None and (r["status"], r["n_rows"], r["sum_total"]) == a for r, a in zip(res, ref))
        # one query at a time, same arrivals
        lat_s = []
        t0 = time.perf_counter()
        for i, d in enumerate(descs):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            ctx.star_join(d)
            lat_s.append((time.perf_counter() - t0 - arrivals[i]) * 1e3)
        wall_s = time.perf_counter() - t0
        e = {"fraction_of_serial_capacity": f, "rate_req_s": lam,
             "served_p50_ms": _pct(lat, 50), "served_p99_ms": _pct(lat, 99),
             "served_rows_per_s": sum(rows) / wall, "launches": st["launches"],
             "mean_requests_per_launch": st["completed"] / max(st["launches"], 1),
             "serial_p50_ms": _pct(lat_s, 50), "serial_p99_ms": _pct(lat_s, 99),
             "serial_rows_per_s": sum(rows) / wall_s, "bit_exact": exact, "checked": len(res),
             "submit_lag_p99_ms": _pct(lag, 99) * 1e3}
        out["rates"].append(e)
        log(f"serve x{f}: served p50 {e['served_p50_ms']:.3f} p99 {e['served_p99_ms']:.3f} ms "
            f"({e['mean_requests_per_launch']:.1f} req/launch) | serial p50 {e['serial_p50_ms']:.3f} "
            f"p99 {e['serial_p99_ms']:.3f} ms")
    for tm in tmpls:
        tm["dim"].release()
    return out


def bench_ssb(ctx, n, reps=3, sf=20):
    """SSB flights 2-4 (src/ssb-jo/q{21..34,41..43}) at n fact rows over SF-sf
    tables: each query in its join orders x {naive, LIP}, and adaptive LIP.
    Every plan's group keys, sums and counts must agree (a payload outside the
    group key is any kept row's value); bytes_alg of a query is the smallest
    over its LIP orders (128 B per touched line, the first column in full; Q4.x
    read the measure lines of both measure columns)."""
    import torch
    import lipjoin as L
    from lipjoin import ssb
    from lipjoin import workloads as W
    # each query gets a context of its own, so its group accumulators are sized
    # for its own group space (the Q3.3 space is 437,500 groups: one copy)
    log(f"ssb flights 2-4: generating {n} lineorder rows at SF {sf}")
    cols = ssb.device_columns(ctx, W.seed_of(7), n, sf)
    torch.cuda.synchronize()
    out = {}
    for name in sorted(ssb.QUERIES):
        q = ssb.QUERIES[name]
        ctx = L.Context(0)
        dims = ssb.dim_descs(q, cols, sf)
        ref = None
        per = {}
        lip_bytes = []
        orders = ssb.orders_of(q)
        for method in ("naive", "lip", "lip-adaptive"):
            for oi, order in (orders.items() if method != "lip-adaptive" else [(0, orders[0])]):
                def call():
                    return ssb.run(ctx, q, cols, n, method, order, sf, dims=dims, groups_cap=1 << 15)
                res = call()
                ref = res["groups"] if ref is None else ref
                assert [(g[0], g[2], g[3]) for g in res["groups"]] == [(g[0], g[2], g[3]) for g in ref], \
                    f"{name} {method} order {oi} disagrees"
                tq, tp = [], []
                for _ in range(reps):
                    r = call()
                    tq.append(r["time_query_ms"])
                    tp.append(r["time_probe_ms"])
                st = res["stats"]
                b = 128 * (sum(st["lines_in"][:len(q.dims) + 1]) + st["sample_lines"])
                if q.measure_sub is not None:
                    b += 128 * st["lines_in"][len(q.dims)]   # the second measure column, same lines
                if method == "lip":
                    lip_bytes.append(b)
                key = method if method == "lip-adaptive" else f"{oi}-{method}"
                per[key] = {"query_ms": statistics.median(tq), "probe_ms": statistics.median(tp)}
        best_naive = min(v["query_ms"] for k, v in per.items() if k.endswith("-naive"))
        ad = per["lip-adaptive"]
        out[name] = {"groups": len(ref), "rows": sum(g[3] for g in ref),
                     "adaptive_query_ms": ad["query_ms"], "adaptive_probe_ms": ad["probe_ms"],
                     "adaptive_rows_per_s": n / (ad["query_ms"] * 1e-3),
                     "bytes_alg": min(lip_bytes),
                     "adaptive_probe_hbm_frac": min(lip_bytes) / (ad["probe_ms"] * 1e-3) / HBM_PEAK,
                     "best_naive_query_ms": best_naive,
                     "worst_naive_query_ms": max(v["query_ms"] for k, v in per.items() if k.endswith("-naive")),
                     "best_lip_query_ms": min(v["query_ms"] for k, v in per.items() if k.endswith("-lip"))}
        ctx.close()
        log(f"ssb {name}: adaptive {ad['query_ms']:.3f} ms (probe {ad['probe_ms']:.3f}, "
            f"frac {out[name]['adaptive_probe_hbm_frac']:.2f}), best naive {best_naive:.3f} ms, "
            f"{len(ref)} groups")
    del cols
    torch.cuda.empty_cache()
    return out


def host_cpu_share():
    """(threads to use, CPUs the OS shows): every CPU this process may run on
    (its affinity set), capped by the cgroup's CPU quota when one is set."""
    visible = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = visible
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n, visible


def cpu_baseline(n_fact, seed, threads):
    """Oracle (C restatement of the reference semantics) on the host cores:
    the LIP plan in join order S->C->D, whole 120M-row input, median of 5
    after one warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from q33_helpers import oracle_q33
    from lipjoin import workloads as W
    host_dims = W.q33_host_dims(seed)
    host_fact = W.q33_host_fact(seed, n_fact, ["lo_suppkey", "lo_custkey", "lo_orderdate", "lo_revenue"])
    ts = []
    res = None
    for i in range(6):
        t0 = time.perf_counter()
        res = oracle_q33(host_dims, host_fact, W.Q33_ORDERS[0], "lip", threads=threads)
        dt = time.perf_counter() - t0
        if i > 0:
            ts.append(dt)
    t = statistics.median(ts)
    return {"value": n_fact / t, "unit": "fact rows/s", "cores": threads, "kind": "port",
            "sample": f"full {n_fact}-row Q3.3 input, LIP order S->C->D (build+probe), median of 5 "
                      f"after 1 warm-up, {threads} threads",
            "seconds": t}, res[0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--fact-rows", type=int, default=120_000_000)
    ap.add_argument("--plan-reps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plans", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=64,
                    help="config-4 batch size (0 skips the batched measurement)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="config-5 requests per arrival rate (0 skips the serving measurement)")
    ap.add_argument("--no-single", action="store_true", help="skip the config 1/2 single-join measurements")
    ap.add_argument("--no-ssb", action="store_true", help="skip the SSB flight 2-4 query measurements")
    ap.add_argument("--no-bf-size", action="store_true",
                    help="skip the reference's logged 2^30-row Bloom-join workload")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}). The star join would shard "
              "by fact rows, with the dimension tables and filters replicated on every GPU and one "
              "int64 all-reduce of the group sums; that path is not built.", file=sys.stderr)
        sys.exit(2)

    import torch
    import lipjoin as L
    from lipjoin import workloads as W

    n = args.fact_rows
    seed = W.seed_of(3)
    ctx = L.Context(0)
    log(f"generating {n} fact rows on the device")
    cols = W.q33_device_columns(ctx, seed, n)
    torch.cuda.synchronize()

    descs = {}

    def run(method, order, bloom_order=None):
        key = (method, tuple(order), tuple(bloom_order) if bloom_order else None)
        if key not in descs:
            descs[key] = W.q33_star_desc(cols, n, method, order, bloom_order)
        return ctx.star_join(descs[key])

    if args.calib_copy_mb:
        # a plain copy of known size: read B bytes, write B bytes (scales the HBM counters)
        words = args.calib_copy_mb * (1 << 20) // 4
        src = torch.ones(words, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: 3 x {args.calib_copy_mb} MiB")
        del src, dst

    # ---- per-plan comparison (naive vs LIP in every join order) -----------------
    plans = {}
    ref_groups = None
    lip_bytes = []
    for name, method, order, bloom_order in W.q33_plans():
        if args.no_plans and name != "lip-adaptive":
            continue
        res = run(method, order, bloom_order)
        if ref_groups is None:
            ref_groups = res["groups"]
        assert res["groups"] == ref_groups, f"plan {name} disagrees"
        ts = []
        tp = []
        for _ in range(args.plan_reps):
            r = run(method, order, bloom_order)
            ts.append(r["time_query_ms"])
            tp.append(r["time_probe_ms"])
        st = res["stats"]
        lines = sum(st["lines_in"][:4]) + st["sample_lines"]
        bytes_alg = 128 * lines
        if method == "lip":
            if bloom_order is None:
                lip_bytes.append(bytes_alg)
        plans[name] = {"query_ms": statistics.median(ts), "probe_ms": statistics.median(tp),
                       "rows_per_s": n / (statistics.median(ts) * 1e-3), "bytes_alg": bytes_alg,
                       "probe_hbm_frac": bytes_alg / (statistics.median(tp) * 1e-3) / HBM_PEAK,
                       "order_used": st["order_used"][:3]}
        log(f"plan {name:18s} query {plans[name]['query_ms']:.3f} ms probe {plans[name]['probe_ms']:.3f} ms "
            f"bytes {bytes_alg / 1e6:.1f} MB")
    speedup = {}
    for i in range(6):
        if f"{i}-naive" in plans and f"{i}-lip" in plans:
            speedup[W.ORDER_NAMES[i]] = plans[f"{i}-naive"]["query_ms"] / plans[f"{i}-lip"]["query_ms"]

    # ---- timed steps: adaptive LIP, whole query per step ----------------------------
    # the timed steps call the C-ABI directly (lj_star_join into a reused result);
    # every step's timings are read back, the groups of the last step are checked
    key = ("lip-adaptive", tuple(W.Q33_ORDERS[0]), None)
    if key not in descs:
        descs[key] = W.q33_star_desc(cols, n, "lip-adaptive", W.Q33_ORDERS[0], None)
    desc = descs[key]
    for _ in range(args.warmup):
        ctx.star_join_raw(desc)
    torch.cuda.synchronize()
    probe_ms = []
    query_ms = []
    raw = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.star_join_raw(desc)
        raw = ctx._raw_res
        probe_ms.append(raw.time_probe_ms)
        query_ms.append(raw.time_query_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    r = ctx.star_join_result()
    if ref_groups is not None:
        assert r["groups"] == ref_groups
    ms_per_step = wall / args.steps * 1e3
    value = n * args.steps / wall

    probe_avg_s = statistics.fmean(probe_ms) * 1e-3
    if lip_bytes:
        bytes_alg = min(lip_bytes)
    else:
        st = r["stats"]
        bytes_alg = 128 * (sum(st["lines_in"][:4]) + st["sample_lines"])
    achieved = bytes_alg / probe_avg_s
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": measured_traffic(),
                "kernel": "lj::probe_kernel (fused LIP scan)", "bytes_alg_per_launch": bytes_alg,
                "probe_ms_avg": probe_avg_s * 1e3}

    out = {
        "metric": METRIC, "value": value, "unit": "fact rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32 keys, int64 sums", "data": "synthetic (seeded, generated on device)",
        "config": {"workload": "SSB Q3.3 star join, LIP with device-chosen filter order",
                   "fact_rows": n, "dims": {"customer": 600000, "supplier": 40000, "date": 2556},
                   "groups": len(ref_groups) if ref_groups is not None else None, "parallelism": "single GPU"},
        "query_ms_median": statistics.median(query_ms),
        "whole_query_hbm_frac": bytes_alg / (statistics.median(query_ms) * 1e-3) / HBM_PEAK,
        "roofline": roofline,
        "lip_vs_naive_speedup": speedup,
        "plans": plans,
    }
    # free the 120M-row columns first
    del cols, descs
    torch.cuda.empty_cache()
    if not args.no_ssb:
        out["ssb_flights"] = bench_ssb(ctx, n)
    if args.batch_instances > 0:
        out["batched"] = bench_batch(ctx, args.batch_instances, reps=5)
    if args.serve_requests > 0:
        out["serving"] = bench_serve(ctx, args.serve_requests)
    if not args.no_single:
        out["single_join"] = bench_single(ctx)
    if not args.no_bf_size:
        out["reference_logged_workload"] = bench_bf_size(ctx)
    if not args.no_cpu_baseline:
        threads, visible = host_cpu_share()
        log(f"cpu baseline on {threads} threads")
        cb, cpu_groups = cpu_baseline(n, seed, threads)
        gpu_groups = [(g, tuple(p), s, c) for (g, p, s, c, _) in r["groups"]]
        cb["matches_gpu"] = cpu_groups == gpu_groups
        cb["host_cpus_visible"] = visible
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
