"""Summaries of rocprofv3 CSV output (kernel statistics and PMC counters).

  python -m rigidmultiblobswall_amd.rocprof_summary stats DIR
      per-kernel calls / total / mean duration (microseconds) from the
      *kernel_stats.csv that `rocprofv3 --kernel-trace --stats
      --output-format csv` wrote under DIR
  python -m rigidmultiblobswall_amd.rocprof_summary traffic --fetch DIR --write DIR
         --kernel SUBSTR --copy-bytes B [--out traffic.json]
      HBM bytes per launch of the kernel whose name contains SUBSTR, from a
      FETCH_SIZE run and a WRITE_SIZE run (`rocprofv3 --pmc ...`), each
      scaled by what the same counter reads for the calibration copy kernel
      (rmw_axpby with beta = 0: B bytes read, B bytes written; bench.py
      --calibrate-copy)
  python -m rigidmultiblobswall_amd.rocprof_summary counters DIR --kernel SUBSTR
      mean per-dispatch value of every counter of a `rocprofv3 --pmc` run for
      the kernels whose name contains SUBSTR
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict
from typing import Dict, List


def _find(d: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[-1]


def kernel_stats(d: str) -> List[Dict]:
    rows = []
    with open(_find(d, "kernel_stats.csv")) as fh:
        for r in csv.DictReader(fh):
            rows.append({"name": r["Name"], "calls": int(r["Calls"]), "total_us": float(r["TotalDurationNs"]) / 1e3,
                         "avg_us": float(r["AverageNs"]) / 1e3, "min_us": float(r.get("MinNs", 0) or 0) / 1e3,
                         "max_us": float(r.get("MaxNs", 0) or 0) / 1e3, "pct": float(r["Percentage"])})
    rows.sort(key=lambda x: -x["total_us"])
    return rows


def counters(d: str) -> Dict[str, Dict[str, List[float]]]:
    """kernel name -> counter name -> per-dispatch values."""
    out: Dict[str, Dict[str, List[float]]] = defaultdict(lambda: defaultdict(list))
    with open(_find(d, "counter_collection.csv")) as fh:
        for r in csv.DictReader(fh):
            out[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def _per_dispatch(cnt, substr: str, counter: str) -> float:
    vals = [v for name, cs in cnt.items() if substr in name for v in cs.get(counter, [])]
    if not vals:
        raise KeyError(f"no {counter} values for a kernel matching {substr!r}")
    return sum(vals) / len(vals)


def _copy_value(cnt, counter: str, substr: str) -> float:
    """Median per-dispatch value of the calibration copy kernel."""
    vals = sorted(v for n, cs in cnt.items() if substr in n for v in cs.get(counter, []))
    if not vals:
        raise KeyError(f"no {counter} values for a copy kernel matching {substr!r}")
    return vals[len(vals) // 2]


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_bytes: float, copy_fetch_dir=None,
            copy_write_dir=None, copy_kernel: str = "axpby_kernel") -> Dict:
    fc = counters(fetch_dir)
    wc = counters(write_dir)
    cf = counters(copy_fetch_dir) if copy_fetch_dir else fc
    cw = counters(copy_write_dir) if copy_write_dir else wc
    fetch_kib = _per_dispatch(fc, kernel, "FETCH_SIZE")
    write_kib = _per_dispatch(wc, kernel, "WRITE_SIZE")
    copy_fetch_kib = _copy_value(cf, "FETCH_SIZE", copy_kernel)
    copy_write_kib = _copy_value(cw, "WRITE_SIZE", copy_kernel)
    f_scale = copy_bytes / (copy_fetch_kib * 1024.0)
    w_scale = copy_bytes / (copy_write_kib * 1024.0)
    tb = fetch_kib * 1024.0 * f_scale + write_kib * 1024.0 * w_scale
    return {"kernel": kernel, "fetch_kib_per_dispatch": fetch_kib, "write_kib_per_dispatch": write_kib,
            "copy_bytes": copy_bytes, "copy_fetch_kib": copy_fetch_kib, "copy_write_kib": copy_write_kib,
            "fetch_scale": f_scale, "write_scale": w_scale, "traffic_bytes": tb}


def counter_summary(d: str, kernel: str) -> Dict[str, float]:
    """Mean per-dispatch value of every counter collected for the kernels
    whose name contains `kernel`."""
    cnt = counters(d)
    out: Dict[str, List[float]] = defaultdict(list)
    for name, cs in cnt.items():
        if kernel in name:
            for c, vals in cs.items():
                out[c].extend(vals)
    if not out:
        raise KeyError(f"no counters for a kernel matching {kernel!r}")
    return {c: sum(v) / len(v) for c, v in sorted(out.items())}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s1 = sub.add_parser("stats")
    s1.add_argument("dir")
    s2 = sub.add_parser("traffic")
    s2.add_argument("--fetch", required=True)
    s2.add_argument("--write", required=True)
    s2.add_argument("--copy-fetch", default=None)
    s2.add_argument("--copy-write", default=None)
    s2.add_argument("--kernel", required=True)
    s2.add_argument("--copy-bytes", type=float, required=True)
    s2.add_argument("--copy-kernel", default="axpby_kernel")
    s2.add_argument("--label", default=None,
                    help="kernel family name recorded in the output (default: the --kernel match)")
    s2.add_argument("--out", default=None)
    s3 = sub.add_parser("counters")
    s3.add_argument("dir")
    s3.add_argument("--kernel", required=True)
    args = ap.parse_args(argv)
    if args.cmd == "counters":
        for c, v in counter_summary(args.dir, args.kernel).items():
            print(f"{c:28s} {v:16.1f}")
        return 0
    if args.cmd == "stats":
        rows = kernel_stats(args.dir)
        print(f"{'kernel':88s} {'calls':>6s} {'total_us':>11s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} "
              f"{'pct':>6s}")
        for r in rows:
            print(f"{r['name'][:88]:88s} {r['calls']:6d} {r['total_us']:11.1f} {r['avg_us']:10.2f} "
                  f"{r['min_us']:10.2f} {r['max_us']:10.2f} {r['pct']:6.2f}")
    else:
        t = traffic(args.fetch, args.write, args.kernel, args.copy_bytes, args.copy_fetch, args.copy_write,
                    args.copy_kernel)
        if args.label:
            t["kernel_match"] = t["kernel"]
            t["kernel"] = args.label
        js = json.dumps(t, indent=1)
        print(js)
        if args.out:
            with open(args.out, "w") as fh:
                fh.write(js + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
