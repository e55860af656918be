"""Summarise rocprofv3 CSV output for the star-join kernels.

  python -m lipjoin.rocprof_summary trace <kernel_trace.csv> [--skip N] [--match S]
  python -m lipjoin.rocprof_summary counters <counter_collection.csv> [--skip N] [--match S]
      [--calib-bytes B]

`trace`: per kernel name, dispatch count and duration statistics (ns) over the
dispatches after the first N of that name (warm-up).  `counters`: per kernel
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counters"])
    ap.add_argument("csv")
    ap.add_argument("--skip", type=int, default=0)
    ap.add_argument("--match", default=None)
    ap.add_argument("--calib-bytes", type=int, default=0)
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.csv, a.skip, a.match)
    else:
        counters(a.csv, a.skip, a.match, a.calib_bytes)


if __name__ == "__main__":
    main()
