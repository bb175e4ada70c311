"""Summarise rocprofv3 CSV output of bench.py into the figures it reports.

Inputs (all written by ``rocprofv3 -f csv``):
  * the kernel statistics of a ``--kernel-trace --stats`` run
    (``*_kernel_stats.csv``: per kernel name, calls and average duration);
  * the per-dispatch counters of a ``--pmc FETCH_SIZE`` run and of a
    ``--pmc WRITE_SIZE`` run (``*_counter_collection.csv``).

FETCH_SIZE and WRITE_SIZE are KiB per dispatch as the TCC sees them.  Each
is scaled by the factor that maps the largest plain copy in the same run
(torch's ``copy_`` of the pristine clip into the work buffer, a
copyBuffer kernel that reads and writes exactly ``--copy-bytes``) to its
known size, so the figures are HBM bytes per launch.

Writes a Markdown summary (the figures and the commands) and merges the
per-launch traffic into a JSON table keyed by workload and kernel, which
bench.py reads for ``roofline.traffic``.

    python -m vbm3d_amd.rocprof_summary --stats S.csv --fetch F.csv --write W.csv \\
        --copy-bytes N --workload "..." --md out.md --json profiles/traffic.json

Without --fetch and --write it tabulates the kernel statistics alone.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re


def short_name(name: str) -> str:
    n = name.replace("void ", "").replace("vb::(anonymous namespace)::", "")
    return n.split("(")[0].strip()


def read_stats(path):
    out = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            out[short_name(row["Name"])] = {"calls": int(row["Calls"]),
                                            "avg_ms": float(row["AverageNs"]) / 1e6,
                                            "total_ms": float(row["TotalDurationNs"]) / 1e6}
    return out


def read_counter(path, counter):
    """[(dispatch id, kernel short name, value KiB)] in dispatch order."""
    rows = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            rows.append((int(row["Dispatch_Id"]), short_name(row["Kernel_Name"]),
                         float(row["Counter_Value"])))
    rows.sort()
    return rows


def calibrate(rows, copy_bytes):
    copies = [v for _, k, v in rows if "copyBuffer" in k]
    if not copies:
        raise SystemExit("no copyBuffer dispatch to calibrate against")
    return copy_bytes / (max(copies) * 1024.0)


def per_kernel(rows, factor):
    agg = collections.OrderedDict()
    for _, k, v in rows:
        a = agg.setdefault(k, [0, 0.0])
        a[0] += 1
        a[1] += v * 1024.0 * factor
    return {k: (n, b / n) for k, (n, b) in agg.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=float)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--md", required=True)
    ap.add_argument("--json")
    ap.add_argument("--commands", nargs="*", default=[])
    ap.add_argument("--min-ms", type=float, default=0.5, help="list kernels above this average")
    args = ap.parse_args()

    stats = read_stats(args.stats)
    if not (args.fetch and args.write):
        lines = [f"# rocprofv3 kernel statistics: {args.workload}", "",
                 "Commands (run from the repository root):", ""]
        lines += [f"    {c}" for c in args.commands]
        lines += ["", "| kernel | calls | avg ms | total ms |", "|---|---|---|---|"]
        for k, s in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"]):
            if s["total_ms"] >= args.min_ms:
                lines.append(f"| `{k}` | {s['calls']} | {s['avg_ms']:.3f} | {s['total_ms']:.1f} |")
        with open(args.md, "w") as f:
            f.write("\n".join(lines) + "\n")
        print("\n".join(lines))
        return
    if args.copy_bytes is None or not args.json:
        raise SystemExit("--fetch and --write need --copy-bytes and --json")
    fe = read_counter(args.fetch, "FETCH_SIZE")
    wr = read_counter(args.write, "WRITE_SIZE")
    ff, fw = calibrate(fe, args.copy_bytes), calibrate(wr, args.copy_bytes)
    kf, kw = per_kernel(fe, ff), per_kernel(wr, fw)

    lines = [f"# rocprofv3 summary: {args.workload}", "",
             "Commands (run from the repository root):", ""]
    lines += [f"    {c}" for c in args.commands]
    lines += ["", f"Calibration: the {args.copy_bytes / 1e6:.1f} MB copyBuffer reads "
                  f"{args.copy_bytes / ff / 1024:.0f} KiB by FETCH_SIZE (factor {ff:.3f}) and writes "
                  f"{args.copy_bytes / fw / 1024:.0f} KiB by WRITE_SIZE (factor {fw:.3f}).", "",
              "| kernel | calls | avg ms | total ms | HBM read MB/launch | HBM write MB/launch |",
              "|---|---|---|---|---|---|"]
    table = {}
    for k, s in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"]):
        if s["avg_ms"] < args.min_ms:
            continue
        rb = kf.get(k, (0, float("nan")))[1]
        wb = kw.get(k, (0, float("nan")))[1]
        lines.append(f"| `{k}` | {s['calls']} | {s['avg_ms']:.3f} | {s['total_ms']:.1f} | "
                     f"{rb / 1e6:.1f} | {wb / 1e6:.1f} |")
        if re.match(r"k_(bm|filter|aggregate|color)", k) and k in kf and k in kw:
            table[k] = {"avg_ms": round(s["avg_ms"], 4), "calls": s["calls"],
                        "fetch_bytes_per_launch": round(rb), "write_bytes_per_launch": round(wb)}
    with open(args.md, "w") as f:
        f.write("\n".join(lines) + "\n")
    tab = {}
    if os.path.exists(args.json):
        with open(args.json) as f:
            tab = json.load(f)
    tab[args.workload] = table
    with open(args.json, "w") as f:
        json.dump(tab, f, indent=1, sort_keys=True)
        f.write("\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
