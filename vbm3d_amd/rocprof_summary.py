"""Summarise rocprofv3 CSV output of bench.py into the figures it reports.

Inputs (all written by ``rocprofv3 -f csv``):
  * the per-dispatch kernel trace of a ``--kernel-trace --stats`` run
    (``*_kernel_trace.csv``: name, start and end of every dispatch);
  * the per-dispatch counters of a ``--pmc FETCH_SIZE`` run and of a
    ``--pmc WRITE_SIZE`` run (``*_counter_collection.csv``);
  * optionally the per-dispatch counters of an instruction / LDS pass
    (``--pmc SQ_INSTS_VALU SQ_INSTS_LDS SQ_WAVES SQ_ACTIVE_INST_VALU
    SQ_THREAD_CYCLES_VALU SQ_LDS_IDX_ACTIVE SQ_LDS_BANK_CONFLICT
    SQ_INSTS_SALU GRBM_GUI_ACTIVE``).

Dispatches are grouped by kernel name; a kernel that runs in both steps of
a run (``--alternate NAME``, e.g. the aggregation, whose template is the
same for both steps) is split by dispatch order into its step-1 and step-2
launches, since every run launches it once per step, step 1 first.
``--runs R`` keeps each kernel's dispatches of the first R runs only
(bench.py: the warm-up and timed device-buffer runs, which precede its
host-buffer leg, whose chunked launches are smaller).

FETCH_SIZE and WRITE_SIZE are KiB per dispatch as the TCC sees them.  Each
is scaled by the factor that maps the largest plain copy in the same run
(torch's ``copy_`` of the pristine clip into the work buffer, a copyBuffer
kernel that reads and writes exactly ``--copy-bytes``) to its known size,
so the figures are HBM bytes per launch.

GRBM_GUI_ACTIVE (the launch's busy cycles) is reported summed over the 8
XCDs, so it is divided by 8 here (rocprof's own formulas take the maximum
over XCDs, which equals the mean for launches that fill every XCD).

LDS figures: LDS busy = SQ_LDS_IDX_ACTIVE / (GRBM_GUI_ACTIVE * CUs), the
fraction of the launch's cycles the CUs' LDS served indexed accesses (its
peak is 128 bytes per clock per CU, 78.6 TB/s over 256 CUs at 2.4 GHz);
bank-conflict share = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE.  VALU busy
= SQ_ACTIVE_INST_VALU / (GRBM_GUI_ACTIVE * CUs) and VALU lane use =
SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU), rocprofv3's own VALUBusy
and VALUUtilization formulas.

Writes a Markdown summary (the figures and the commands) and merges the
per-launch traffic into a JSON table keyed by workload and kernel, which
bench.py reads for ``roofline.traffic``.

    python -m vbm3d_amd.rocprof_summary --trace T.csv --fetch F.csv --write W.csv \\
        [--insts I.csv] --copy-bytes N --workload "..." --md out.md --json profiles/traffic.json

``--stats S.csv`` (the ``*_kernel_stats.csv`` of the same run: calls and
average duration per kernel name) may stand in for ``--trace`` when no
kernel needs splitting; without --fetch and --write the summary tabulates
the kernel durations alone.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re

CUS = 256
XCDS = 8  # GRBM_GUI_ACTIVE arrives summed over the XCDs; its per-XCD mean stands in for rocprof's max
LDS_PEAK_TBS = 128 * CUS * 2.4e9 / 1e12


def short_name(name: str) -> str:
    n = name.replace("void ", "").replace("vb::(anonymous namespace)::", "")
    return n.split("(")[0].strip()


RUNS = None  # keep only each kernel's dispatches of the first RUNS runs


def label(names_in_order, alternate):
    """Per dispatch (in dispatch order): the kernel name, with ' (step 1)'
    / ' (step 2)' appended alternately for the kernels in `alternate`, or
    None for a dispatch past the first RUNS runs of its kernel."""
    seen = collections.Counter()
    out = []
    for n in names_in_order:
        per_run = 2 if n in alternate else 1
        if RUNS is not None and "copyBuffer" not in n and seen[n] >= RUNS * per_run:
            out.append(None)
        elif n in alternate:
            out.append(f"{n} (step {1 + seen[n] % 2})")
        else:
            out.append(n)
        seen[n] += 1
    return out


def read_trace(path, alternate):
    rows = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            rows.append((int(row["Dispatch_Id"]), short_name(row["Kernel_Name"]),
                         (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e6))
    rows.sort()
    labels = label([k for _, k, _ in rows], alternate)
    out = collections.OrderedDict()
    for (_, _, ms), lab in zip(rows, labels):
        if lab is None:
            continue
        e = out.setdefault(lab, [0, 0.0])
        e[0] += 1
        e[1] += ms
    return {k: {"calls": n, "avg_ms": t / n, "total_ms": t} for k, (n, t) in out.items()}


def read_stats(path):
    out = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            out[short_name(row["Name"])] = {"calls": int(row["Calls"]),
                                            "avg_ms": float(row["AverageNs"]) / 1e6,
                                            "total_ms": float(row["TotalDurationNs"]) / 1e6}
    return out


def read_counters(path, alternate):
    """{label: {counter: [values per dispatch]}} in dispatch order."""
    per = collections.defaultdict(dict)
    names = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            d = int(row["Dispatch_Id"])
            names[d] = short_name(row["Kernel_Name"])
            per[d][row["Counter_Name"]] = float(row["Counter_Value"])
    order = sorted(per)
    labels = label([names[d] for d in order], alternate)
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, lab in zip(order, labels):
        if lab is None:
            continue
        for c, v in per[d].items():
            out[lab][c].append(v)
    return out


def mean(v):
    return sum(v) / len(v) if v else float("nan")


def main():
    ap = argparse.ArgumentParser()
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--trace")
    src.add_argument("--stats")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--insts")
    ap.add_argument("--copy-bytes", type=float)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--md", required=True)
    ap.add_argument("--json")
    ap.add_argument("--alternate", nargs="*", default=[])
    ap.add_argument("--runs", type=int, help="only the first RUNS runs (dispatches of each kernel in order); "
                                             "bench.py's device-buffer runs come before its host-buffer leg")
    ap.add_argument("--commands", nargs="*", default=[])
    ap.add_argument("--min-ms", type=float, default=0.5, help="list kernels whose average (total, without counters) reaches this")
    args = ap.parse_args()
    alt = set(args.alternate)
    global RUNS
    RUNS = args.runs

    stats = read_trace(args.trace, alt) if args.trace else read_stats(args.stats)
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
    if args.copy_bytes is None or args.json is None:
        raise SystemExit("--fetch/--write need --copy-bytes and --json")
    fe = read_counters(args.fetch, alt)
    wr = read_counters(args.write, alt)
    copies = [v for k, c in fe.items() if "copyBuffer" in k for v in c.get("FETCH_SIZE", [])]
    copies_w = [v for k, c in wr.items() if "copyBuffer" in k for v in c.get("WRITE_SIZE", [])]
    if not copies or not copies_w:
        raise SystemExit("no copyBuffer dispatch to calibrate against")
    ff = args.copy_bytes / (max(copies) * 1024.0)
    fw = args.copy_bytes / (max(copies_w) * 1024.0)
    ins = read_counters(args.insts, alt) if args.insts else {}

    lines = [f"# rocprofv3 summary: {args.workload}", "",
             "Commands (run from the repository root):", ""]
    lines += [f"    {c}" for c in args.commands]
    lines += ["", f"Calibration: the {args.copy_bytes / 1e6:.1f} MB copyBuffer reads "
                  f"{args.copy_bytes / ff / 1024:.0f} KiB by FETCH_SIZE (factor {ff:.3f}) and writes "
                  f"{args.copy_bytes / fw / 1024:.0f} KiB by WRITE_SIZE (factor {fw:.3f}).", "",
              "Kernels whose template serves both steps are split by dispatch order (step 1 first).", "",
              "| kernel | calls | avg ms | HBM read MB/launch | HBM write MB/launch | HBM GB/s |",
              "|---|---|---|---|---|---|"]
    table = {}
    for k, s in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"]):
        if s["avg_ms"] < args.min_ms:
            continue
        rb = mean(fe.get(k, {}).get("FETCH_SIZE", [])) * 1024.0 * ff
        wb = mean(wr.get(k, {}).get("WRITE_SIZE", [])) * 1024.0 * fw
        lines.append(f"| `{k}` | {s['calls']} | {s['avg_ms']:.3f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} | "
                     f"{(rb + wb) / (s['avg_ms'] * 1e-3) / 1e9:.0f} |")
        if re.match(r"k_(bm|filter|aggregate|color)", k):
            table[k] = {"avg_ms": round(s["avg_ms"], 4), "calls": s["calls"],
                        "fetch_bytes_per_launch": round(rb), "write_bytes_per_launch": round(wb)}
    if ins:
        lines += ["", f"Instructions and LDS (one wavefront instruction counts once; LDS peak "
                      f"{LDS_PEAK_TBS:.1f} TB/s = 128 B/clk/CU):", "",
                  "| kernel | waves | VALU instr/wave | LDS instr/wave | SALU instr/wave | VALU busy | "
                  "VALU lane use | LDS busy | LDS bank-conflict share |",
                  "|---|---|---|---|---|---|---|---|---|"]
        for k, s in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"]):
            if s["avg_ms"] < args.min_ms or k not in ins:
                continue
            c = ins[k]
            waves = mean(c.get("SQ_WAVES", []))
            gui = mean(c.get("GRBM_GUI_ACTIVE", [])) / XCDS
            act = mean(c.get("SQ_ACTIVE_INST_VALU", []))
            thr = mean(c.get("SQ_THREAD_CYCLES_VALU", []))
            lds = mean(c.get("SQ_LDS_IDX_ACTIVE", []))
            bank = mean(c.get("SQ_LDS_BANK_CONFLICT", []))
            lines.append(
                f"| `{k}` | {waves:.0f} | {mean(c.get('SQ_INSTS_VALU', [])) / waves:.0f} | "
                f"{mean(c.get('SQ_INSTS_LDS', [])) / waves:.0f} | {mean(c.get('SQ_INSTS_SALU', [])) / waves:.0f} | "
                f"{100 * act / (gui * CUS):.1f} % | {100 * thr / (64 * act):.1f} % | "
                f"{100 * lds / (gui * CUS):.1f} % | {100 * bank / lds if lds else 0:.1f} % |")
            if k in table:
                table[k].update({"valu_busy": round(act / (gui * CUS), 4), "lds_busy": round(lds / (gui * CUS), 4),
                                 "lds_bank_conflict_share": round(bank / lds, 4) if lds else 0.0})
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
