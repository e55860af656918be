"""Summarise rocprofv3 CSV output for the star-join kernels.

  python -m lipjoin.rocprof_summary trace <kernel_trace.csv> [--skip N] [--match S]
  python -m lipjoin.rocprof_summary counters <counter_collection.csv> [--skip N] [--match S]
      [--calib-bytes B]

  python -m lipjoin.rocprof_summary workloads --trace <kernel_trace.csv>
      [--counters <counter_collection.csv> ...] --calib-bytes B --reps R WORKLOAD [WORKLOAD ...]

`trace`: per kernel name, dispatch count and duration statistics (ns) over the
dispatches after the first N of that name (warm-up).  `workloads`: the probe
dispatches of a `lipjoin.profile_runs` run (R per workload, in the order the
workloads were named), one line per workload: duration without the first
(warm-up) dispatch, and each counter's mean (FETCH_SIZE / WRITE_SIZE scaled by
the copy calibration, TCC hit rate, LDS bank-conflict cycles per LDS
instruction).  `counters`: per kernel
name, mean counter value per dispatch; with --calib-bytes, the counter read by
the largest `__amd_rocclr_copyBuffer` dispatch is divided into B to give the
counter's scale factor (a plain copy of B bytes reads B and writes B bytes).
"""
import argparse
import csv
import statistics
import sys
from collections import OrderedDict


def _short(name, width=72):
    return name if len(name) <= width else name[: width - 3] + "..."


def trace(path, skip=0, match=None, out=None):
    out = out or sys.stdout
    per = OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if match and match not in name:
                continue
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            per.setdefault(name, []).append(d)
    print(f"{'kernel':74s} {'calls':>6s} {'avg_ns':>11s} {'median_ns':>11s} {'min_ns':>10s} {'max_ns':>10s}", file=out)
    res = {}
    for name, ds in per.items():
        ds = ds[skip:] if len(ds) > skip else ds
        res[name] = {"calls": len(ds), "avg_ns": statistics.fmean(ds), "median_ns": statistics.median(ds),
                     "min_ns": min(ds), "max_ns": max(ds)}
        v = res[name]
        print(f"{_short(name):74s} {v['calls']:6d} {v['avg_ns']:11.1f} {v['median_ns']:11.1f} "
              f"{v['min_ns']:10d} {v['max_ns']:10d}", file=out)
    return res


def counters(path, skip=0, match=None, calib_bytes=0, out=None):
    out = out or sys.stdout
    per = OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            name, ctr = r["Kernel_Name"], r["Counter_Name"]
            key = (name, ctr)
            per.setdefault(key, []).append(float(r["Counter_Value"]))
    scale = {}
    if calib_bytes:
        for (name, ctr), vs in per.items():
            if name == "__amd_rocclr_copyBuffer":
                top = max(vs)
                if top * 1024 > calib_bytes / 8:  # the big copy, not small result copies
                    scale[ctr] = calib_bytes / (top * 1024.0)
    print(f"{'kernel':74s} {'counter':>11s} {'calls':>6s} {'mean_KiB':>14s} {'scaled_bytes':>16s}", file=out)
    res = {}
    for (name, ctr), vs in per.items():
        if match and match not in name:
            continue
        vs = vs[skip:] if len(vs) > skip else vs
        m = statistics.fmean(vs)
        sb = m * 1024.0 * scale.get(ctr, 1.0)
        res.setdefault(name, {})[ctr] = {"calls": len(vs), "mean_kib": m, "scaled_bytes": sb}
        print(f"{_short(name):74s} {ctr:>11s} {len(vs):6d} {m:14.3f} {sb:16.0f}", file=out)
    for ctr, f in scale.items():
        print(f"scale[{ctr}] = {f:.5f} (copy of {calib_bytes} bytes)", file=out)
    return res, scale


def _probe_rows(path, key):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if "probe_kernel" in r["Kernel_Name"] or "probe_multi_kernel" in r["Kernel_Name"]:
                rows.append(r)
    return rows


def _split(vals, names, reps):
    if len(vals) != len(names) * reps:
        raise SystemExit(f"{len(vals)} probe dispatches for {len(names)} workloads x {reps} reps")
    return {n: vals[i * reps:(i + 1) * reps] for i, n in enumerate(names)}


def workloads(trace_csv, counter_csvs, names, reps, calib_bytes, out=None):
    out = out or sys.stdout
    dur = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in _probe_rows(trace_csv, None)]
    kern = [r["Kernel_Name"] for r in _probe_rows(trace_csv, None)]
    per = {n: {"avg_us": statistics.fmean(v[1:] if reps > 1 else v) / 1e3} for n, v in _split(dur, names, reps).items()}
    for n, k in _split(kern, names, reps).items():
        per[n]["kernel"] = k[0].replace("void lj::", "").split("(")[0]
    scale = {}
    for path in counter_csvs:
        ctr_vals = {}
        with open(path) as f:
            rows = list(csv.DictReader(f))
        for r in rows:
            if r["Kernel_Name"] == "__amd_rocclr_copyBuffer" and calib_bytes and \
                    r["Counter_Name"] in ("FETCH_SIZE", "WRITE_SIZE"):
                top = float(r["Counter_Value"])
                if top * 1024 > calib_bytes / 8:
                    scale[r["Counter_Name"]] = calib_bytes / (top * 1024.0)
            elif "probe_kernel" in r["Kernel_Name"] or "probe_multi_kernel" in r["Kernel_Name"]:
                ctr_vals.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        for ctr, vals in ctr_vals.items():
            for n, v in _split(vals, names, reps).items():
                per[n][ctr] = statistics.fmean(v[1:] if reps > 1 else v)
    cols = ["kernel", "avg_us"]
    extra = []
    for n, d in per.items():
        if "FETCH_SIZE" in d:
            d["fetch_MB"] = d["FETCH_SIZE"] * 1024 * scale.get("FETCH_SIZE", 1.0) / 1e6
        if "WRITE_SIZE" in d:
            d["write_MB"] = d["WRITE_SIZE"] * 1024 * scale.get("WRITE_SIZE", 1.0) / 1e6
        if "fetch_MB" in d:
            d["fetch_TBps"] = d["fetch_MB"] / d["avg_us"]   # MB per us = TB/s
        if "TCC_HIT_sum" in d and "TCC_MISS_sum" in d:
            d["tcc_hit"] = d["TCC_HIT_sum"] / max(1.0, d["TCC_HIT_sum"] + d["TCC_MISS_sum"])
        if "SQ_LDS_BANK_CONFLICT" in d and "SQ_INSTS_LDS" in d:
            d["ldsconf_per_inst"] = d["SQ_LDS_BANK_CONFLICT"] / max(1.0, d["SQ_INSTS_LDS"])
    for k in ("fetch_MB", "write_MB", "fetch_TBps", "TCC_HIT_sum", "TCC_MISS_sum", "tcc_hit", "SQ_INSTS_LDS",
              "SQ_LDS_BANK_CONFLICT", "ldsconf_per_inst", "SQ_INSTS_VALU", "SQ_WAVES"):
        if any(k in d for d in per.values()):
            extra.append(k)
    hdr = "| workload | " + " | ".join(cols + extra) + " |"
    print(hdr, file=out)
    print("|" + "---|" * (len(cols) + len(extra) + 1), file=out)
    for n, d in per.items():
        cells = [n]
        for k in cols + extra:
            v = d.get(k, "")
            if isinstance(v, float):
                v = f"{v:.3g}" if abs(v) < 1e4 else f"{v:.4g}"
            cells.append(str(v))
        print("| " + " | ".join(cells) + " |", file=out)
    for ctr, f in scale.items():
        print(f"\nscale[{ctr}] = {f:.5f} (copy of {calib_bytes} bytes)", file=out)
    return per


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counters", "workloads"])
    ap.add_argument("csv", nargs="*")
    ap.add_argument("--skip", type=int, default=0)
    ap.add_argument("--match", default=None)
    ap.add_argument("--calib-bytes", type=int, default=0)
    ap.add_argument("--trace")
    ap.add_argument("--counters", nargs="*", default=[])
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    if a.mode == "workloads":
        workloads(a.trace, a.counters, a.csv, a.reps, a.calib_bytes)
    elif a.mode == "trace":
        a.csv = a.csv[0]
        trace(a.csv, a.skip, a.match)
    else:
        counters(a.csv[0], a.skip, a.match, a.calib_bytes)


if __name__ == "__main__":
    main()
