"""Summarises a rocprofv3 kernel trace (``--kernel-trace --output-format csv``).

Krylov kernels that are launched after the solve has converged return at once
(they read the device-side stop flag); those launches are listed separately so
the average duration of the launches that did work can be compared with the
library's own HIP-event timing.

usage: python -m shiokaze_amd.rocprof_summary <prefix>_kernel_trace.csv [--noop-us 5] [--command "..."]
       [--by-grid]
"""
from __future__ import annotations

import argparse
import csv
import re
import statistics


def short_name(name: str) -> str:
    m = re.search(r"(k_[A-Za-z0-9_]+)(<[^()]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name.split("(")[0][:60]


def summarise(path: str, noop_us: float, by_grid: bool = False) -> list[dict]:
    """Launch statistics per kernel, or per (kernel, grid size) with by_grid (one
    convolution kernel serves several layers; the grid tells them apart)."""
    by: dict[str, list[float]] = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            key = short_name(r["Kernel_Name"])
            if by_grid:
                key += " grid " + "x".join(r.get(c, "?") for c in ("Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"))
            by.setdefault(key, []).append(d)
    out = []
    for k, v in by.items():
        work = [x for x in v if x >= noop_us]
        out.append({"kernel": k, "calls": len(v), "work_calls": len(work),
                    "avg_all_us": statistics.mean(v), "avg_work_us": statistics.mean(work) if work else 0.0,
                    "median_work_us": statistics.median(work) if work else 0.0, "total_ms": sum(v) / 1e3})
    out.sort(key=lambda r: -r["total_ms"])
    return out


def sequence(path: str, start: str, length: int) -> list[tuple[str, str, float, int]]:
    """Median duration of each launch position in a repeated launch sequence: the trace in
    start-time order is cut at every launch of kernel `start`, and position i of each cut
    (up to `length`) is one sample of (kernel, grid) at i -- the layers of one network
    application, say, which share kernels but not positions."""
    with open(path) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    runs: list[list[tuple[str, str, float]]] = []
    for r in rows:
        name = short_name(r["Kernel_Name"])
        grid = "x".join(r.get(c, "?") for c in ("Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"))
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        if name == start:
            runs.append([])
        if runs:
            runs[-1].append((name, grid, d))
    full = [run for run in runs if len(run) >= length]
    out = []
    for i in range(length):
        keys = {(run[i][0], run[i][1]) for run in full}
        if len(keys) != 1:
            break  # the sequence does not repeat beyond here
        name, grid = keys.pop()
        out.append((name, grid, statistics.median(run[i][2] for run in full), len(full)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--noop-us", type=float, default=5.0)
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--by-grid", action="store_true", help="one row per (kernel, grid size)")
    ap.add_argument("--sequence", default=None, metavar="KERNEL",
                    help="median duration per position of the launch sequence that starts at KERNEL")
    ap.add_argument("--length", type=int, default=20, help="positions of --sequence")
    a = ap.parse_args()
    if a.sequence:
        seq = sequence(a.trace, a.sequence, a.length)
        print(f"launch sequence from `{a.sequence}`, median over {seq[0][3] if seq else 0} repetitions\n")
        print("| position | kernel | grid | median (us) |")
        print("|---:|---|---|---:|")
        for i, (name, grid, med, _) in enumerate(seq):
            print(f"| {i} | `{name}` | {grid} | {med:.2f} |")
        return
    rows = summarise(a.trace, a.noop_us, a.by_grid)
    if a.command:
        print(f"command: `{a.command}`\n")
    print(f"launches shorter than {a.noop_us:g} us are stop-flag early exits\n")
    print("| kernel | calls | calls that did work | avg of those (us) | median (us) | avg all (us) | total (ms) |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for r in rows[:a.top]:
        print(f"| `{r['kernel']}` | {r['calls']} | {r['work_calls']} | {r['avg_work_us']:.2f} | "
              f"{r['median_work_us']:.2f} | {r['avg_all_us']:.2f} | {r['total_ms']:.2f} |")


if __name__ == "__main__":
    main()
