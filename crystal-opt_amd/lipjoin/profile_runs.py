"""Run named workloads of the star-join path a few times, for a profiler.

  python -m lipjoin.profile_runs WORKLOAD [WORKLOAD ...] [--reps N] [--calib-copy-mb M]

Workloads (synthetic, seeded, generated on the device):
  q33:<plan>            SSB Q3.3 at 120,000,000 fact rows, plan as in the reference's
                        binaries (0-lip ... 5-naive, 3-lip-best-bloom) or lip-adaptive
  ssb:<query>:<plan>    another SSB query (q21 ... q43) at 120,000,000 rows over SF-20 tables
  config1               the plain join: 2^20-row dimension, 2^24 fact rows, naive
  config2:<sel>:<bloom> the Bloom join at 2^28 rows, selectivity <sel> %, bloom "tuned",
                        "semi" (SUM(measure) only, filter tuned for it), "none" (naive)
                        or "mult-<lg>-<k>" / "mod-<lg>"
  batch                 the 64-instance batch (one graph replay per repetition)
--calib-copy-mb M first runs a plain device copy of M MiB (a reference for the HBM
counters: it reads and writes M MiB).  Each workload line printed to stderr names
what runs, so a profiler's per-kernel output can be matched to it.
"""
import argparse
import sys


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _q33(ctx, plan, reps, cols_cache):
    from . import workloads as W
    import torch
    if "q33" not in cols_cache:
        cols_cache["q33"] = W.q33_device_columns(ctx, W.seed_of(3), W.Q33_FACT_ROWS)
        torch.cuda.synchronize()
    cols = cols_cache["q33"]
    name, method, order, bo = [p for p in W.q33_plans() if p[0] == plan][0]
    desc = W.q33_star_desc(cols, W.Q33_FACT_ROWS, method, order, bo)
    for _ in range(reps):
        r = ctx.star_join(desc)
    return f"q33 {plan}: probe {r['time_probe_ms']:.3f} ms, lines {r['stats']['lines_in'][:4]}"


def _ssb(ctx, query, plan, reps, cols_cache):
    import torch
    from . import Context, ssb
    from . import workloads as W
    n, sf = 120_000_000, 20
    if "ssb" not in cols_cache:
        cols_cache["ssb"] = ssb.device_columns(ctx, W.seed_of(7), n, sf)
        torch.cuda.synchronize()
    cols = cols_cache["ssb"]
    q = ssb.QUERIES[query]
    method, order = ssb.parse_plan(q, plan)
    dims = ssb.dim_descs(q, cols, sf)
    with Context(0) as c2:          # its own group accumulators, as the bench runs it
        for _ in range(reps):
            r = ssb.run(c2, q, cols, n, method, order, sf, dims=dims, groups_cap=1 << 15)
    return f"ssb {query} {plan}: probe {r['time_probe_ms']:.3f} ms, lines {r['stats']['lines_in'][:5]}"


def _config1(ctx, reps):
    from . import dim_desc
    from . import workloads as W
    seed, nd, nf = W.seed_of(1), 1 << 20, 1 << 24
    dk, da, dfk, dm = W.single_join_device(ctx, seed, nd, nf, 99)
    desc = dim_desc(dk, nd, nd, attr=da, predicate=W.single_join_pred(10))
    for _ in range(reps):
        r = ctx.hash_join(dfk, dm, nf, desc, method="naive")
    return f"config1: probe {r['time_probe']:.3f} ms, matches {r['n_match']}"


def _config2(ctx, sel, bloom_s, reps):
    from . import dim_desc
    from . import workloads as W
    seed, nd, nf = W.seed_of(2), 1 << 22, 1 << 28
    dk, da, dfk, dm = W.single_join_device(ctx, seed, nd, nf, 9999)
    thr = 100 * int(sel)
    if bloom_s in ("tuned", "semi"):
        bloom, _ = W.tuned_bloom(ctx, {"n_dim": nd, "threshold": thr}, semi_join=bloom_s == "semi")
    elif bloom_s == "none":
        bloom = None
    else:
        parts = bloom_s.split("-")
        bloom = ("mod", 1 << int(parts[1])) if parts[0] == "mod" else ("mult", 1 << int(parts[1]), int(parts[2]), 0)
    desc = dim_desc(dk, nd, nd, attr=da, predicate=W.single_join_pred(thr), bloom=bloom)
    for _ in range(reps):
        r = ctx.hash_join(dfk, dm, nf, desc, method="naive" if bloom is None else "lip",
                          sum_only=bloom_s == "semi")
    return f"config2 s={sel}% {bloom_s} {bloom}: probe {r['time_probe']:.3f} ms, matches {r['n_match']}"


def _batch(ctx, reps):
    import torch
    from . import workloads as W
    specs = W.batch_specs(W.seed_of(4))
    cols, descs = [], []
    for sp in specs:
        c = W.batch_instance_device(ctx, sp)
        bloom, _ = W.tuned_bloom(ctx, sp, semi_join=True)
        cols.append(c)
        descs.append(W.batch_desc(c, sp, method="lip", bloom=bloom))
    torch.cuda.synchronize()
    b = ctx.batch(descs, use_graph=True)
    for _ in range(reps):
        res = b.run()
    b.close()
    return f"batch: {len(res)} instances, {sum(r['n_rows'] for r in res)} rows kept"


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m lipjoin.profile_runs", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workloads", nargs="+")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--calib-copy-mb", type=int, default=0)
    args = ap.parse_args(argv)
    import torch
    from . import Context
    ctx = Context(0)
    if args.calib_copy_mb:
        words = args.calib_copy_mb * (1 << 20) // 4
        src = torch.ones(words, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        _log(f"calibration copy: 3 x {args.calib_copy_mb} MiB")
        del src, dst
    cache = {}
    for w in args.workloads:
        parts = w.split(":")
        if parts[0] == "q33":
            msg = _q33(ctx, parts[1], args.reps, cache)
        elif parts[0] == "ssb":
            msg = _ssb(ctx, parts[1], parts[2], args.reps, cache)
        elif parts[0] == "config1":
            msg = _config1(ctx, args.reps)
        elif parts[0] == "config2":
            msg = _config2(ctx, parts[1], parts[2], args.reps)
        elif parts[0] == "batch":
            msg = _batch(ctx, args.reps)
        else:
            raise SystemExit(f"unknown workload {w!r}")
        torch.cuda.synchronize()
        _log(msg)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
