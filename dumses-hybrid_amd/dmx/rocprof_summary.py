"""Summaries of rocprofv3 CSV output for the MHD step kernels.

    python -m dmx.rocprof_summary stats  <kernel_stats.csv>
    python -m dmx.rocprof_summary trace  <kernel_trace.csv> [--last N]
    python -m dmx.rocprof_summary pmc    <counter_collection.csv> [...] [--last N]

`stats` prints the kernel-trace statistics table; `trace` prints the same
from the per-dispatch trace; `pmc` prints, per kernel, the mean of every
collected counter per dispatch, plus the HBM byte calibration rows (torch
fill / elementwise kernels of known size) when the run included them.
With --last N only the last N dispatches of each kernel count: the
launch-shape autotuner runs every candidate before the timed steps, so the
last dispatches are the steady-state ones of the chosen shape.
"""
from __future__ import annotations

import collections
import csv
import sys


def short(name: str) -> str:
    name = name.replace("void ", "")
    if name.startswith("at::native::"):
        if "FillFunctor" in name:
            return "calib: torch fill 1 GiB (write)"
        if "Mul" in name or "mul" in name or "BinaryFunctor" in name or "AUnaryFunctor" in name:
            return "calib: torch mul 1 GiB read + 1 GiB write"
        return name[:60]
    return name.split("(")[0]


def stats(path: str) -> str:
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | mean us | min us | max us | % |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def trace(path: str, last: int | None = None) -> str:
    runs = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        runs[short(r["Kernel_Name"])].append((int(r["Dispatch_Id"]),
                                              (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    rows = []
    for k, v in runs.items():
        v.sort()
        d = [t for _, t in (v[-last:] if last else v)]
        rows.append((sum(d), k, len(v), len(d), sum(d) / len(d), min(d), max(d)))
    out = ["| kernel | calls | counted | mean us | min us | max us |", "|---|---:|---:|---:|---:|---:|"]
    for _, k, n, c, mean, lo, hi in sorted(rows, reverse=True):
        out.append(f"| `{k}` | {n} | {c} | {mean:.1f} | {lo:.1f} | {hi:.1f} |")
    return "\n".join(out)


def pmc(paths: list[str], last: int | None = None) -> str:
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    meta = {}
    for path in paths:
        rows = list(csv.DictReader(open(path)))
        if last:
            ids = collections.defaultdict(set)
            for r in rows:
                ids[r["Kernel_Name"]].add(int(r["Dispatch_Id"]))
            keep = {k: set(sorted(v)[-last:]) for k, v in ids.items()}
            rows = [r for r in rows if int(r["Dispatch_Id"]) in keep[r["Kernel_Name"]]]
        for r in rows:
            k = short(r["Kernel_Name"])
            if k.startswith("__amd_rocclr"):
                continue
            vals[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            meta[k] = (r["Grid_Size"], r["Workgroup_Size"], r["LDS_Block_Size"], r["VGPR_Count"],
                       r["SGPR_Count"], r["Scratch_Size"])
    counters = sorted({c for k in vals for c in vals[k]})
    head = "| kernel | grid | wg | LDS B | VGPR | SGPR | scratch | " + " | ".join(counters) + " |"
    out = [head, "|" + "---|" * (7 + len(counters))]
    for k in sorted(vals):
        g, w, lds, vg, sg, scr = meta[k]
        cells = []
        for c in counters:
            v = vals[k].get(c)
            cells.append(f"{sum(v) / len(v):.6g}" if v else "")
        out.append(f"| `{k}` | {g} | {w} | {lds} | {vg} | {sg} | {scr} | " + " | ".join(cells) + " |")
    return "\n".join(out)


def main(argv):
    if len(argv) < 3 or argv[1] not in ("stats", "trace", "pmc"):
        print(__doc__)
        return 2
    args, last = list(argv[2:]), None
    if "--last" in args:
        i = args.index("--last")
        last = int(args[i + 1])
        del args[i:i + 2]
    if argv[1] == "stats":
        print(stats(args[0]))
    elif argv[1] == "trace":
        print(trace(args[0], last))
    else:
        print(pmc(args, last))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
