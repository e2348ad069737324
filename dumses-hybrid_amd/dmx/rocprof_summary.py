"""Summaries of rocprofv3 CSV output for the MHD step kernels.

    python -m dmx.rocprof_summary stats  <kernel_stats.csv>
    python -m dmx.rocprof_summary pmc    <counter_collection.csv> [...]

`stats` prints the kernel-trace statistics table; `pmc` prints, per kernel,
the mean of every collected counter per dispatch, plus the HBM byte
calibration rows (torch fill / elementwise kernels of known size) when the
run included them.
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


def pmc(paths: list[str]) -> str:
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    meta = {}
    for path in paths:
        for r in csv.DictReader(open(path)):
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
    if len(argv) < 3 or argv[1] not in ("stats", "pmc"):
        print(__doc__)
        return 2
    print(stats(argv[2]) if argv[1] == "stats" else pmc(argv[2:]))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
