"""Summarises a rocprofv3 kernel trace (``--kernel-trace --output-format csv``).

Krylov kernels that are launched after the solve has converged return at once
(they read the device-side stop flag); those launches are listed separately so
the average duration of the launches that did work can be compared with the
library's own HIP-event timing.

usage: python -m shiokaze_amd.rocprof_summary <prefix>_kernel_trace.csv [--noop-us 5] [--command "..."]
"""
from __future__ import annotations

import argparse
import csv
import re
import statistics


def short_name(name: str) -> str:
    m = re.search(r"(k_[A-Za-z0-9_]+)(<[^()]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name.split("(")[0][:60]


def summarise(path: str, noop_us: float) -> list[dict]:
    by: dict[str, list[float]] = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            by.setdefault(short_name(r["Kernel_Name"]), []).append(d)
    out = []
    for k, v in by.items():
        work = [x for x in v if x >= noop_us]
        out.append({"kernel": k, "calls": len(v), "work_calls": len(work),
                    "avg_all_us": statistics.mean(v), "avg_work_us": statistics.mean(work) if work else 0.0,
                    "median_work_us": statistics.median(work) if work else 0.0, "total_ms": sum(v) / 1e3})
    out.sort(key=lambda r: -r["total_ms"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--noop-us", type=float, default=5.0)
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=25)
    a = ap.parse_args()
    rows = summarise(a.trace, a.noop_us)
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
