"""Per-(kernel, grid) summary of a rocprofv3 kernel trace.

rocprofv3 --stats aggregates by kernel name, which merges the launches of one
kernel on every multigrid level; this splits them by grid size so a kernel's
launches at one problem size can be compared with bench.py's HIP-event time.

    python -m vox8.trace_summary <..._kernel_trace.csv> [top_n] [name_filter]

prints a markdown table: total ms, calls, mean/min/max us, grid, kernel.
"""
from __future__ import annotations

import csv
import sys
from collections import defaultdict


def summarize(path: str, top: int = 40, name_filter: str = "") -> str:
    agg = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if name_filter and name_filter not in name:
                continue
            grid = r.get("Grid_Size") or r.get("Grid_Size_X") or ""
            agg[(name, grid)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    total = sum(sum(v) for v in agg.values())
    rows = sorted(agg.items(), key=lambda kv: -sum(kv[1]))
    out = [f"Total GPU kernel time: {total / 1e3:.1f} ms over {sum(len(v) for v in agg.values())} launches.",
           "", "| total ms | calls | mean us | min us | max us | grid | kernel |",
           "|---:|---:|---:|---:|---:|---:|---|"]
    for (name, grid), t in rows[:top]:
        short = name.split("(")[0].replace("|", "\\|")
        out.append(f"| {sum(t) / 1e3:.2f} | {len(t)} | {sum(t) / len(t):.1f} | {min(t):.1f} | {max(t):.1f} "
                   f"| {grid} | `{short}` |")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    a = sys.argv[1:]
    if not a:
        print(__doc__)
        sys.exit(2)
    print(summarize(a[0], int(a[1]) if len(a) > 1 else 40, a[2] if len(a) > 2 else ""), end="")
