"""Summaries of rocprofv3 CSV output: kernel durations and PMC counters per
kernel and grid size (rocprofv3's own --stats table groups by name only, which
mixes launches of different shapes -- e.g. the batched run's per-stream
histogram groups and a stage launched over all histograms).

python -m wassersteindictionarylearning_amd._profile trace <..._kernel_trace.csv> [--top N]
python -m wassersteindictionarylearning_amd._profile pmc <..._counter_collection.csv> [--top N]

Prints Markdown tables: calls, average / min / max duration in microseconds
(trace), or the average counter value per dispatch in the counter's unit
(pmc; FETCH_SIZE and WRITE_SIZE are KiB).
"""
from __future__ import annotations

import argparse
import csv
import re
import sys
from collections import defaultdict


def _short(name: str) -> str:
    name = re.sub(r"\(.*\)$", "", name.strip())
    name = name.replace("void ", "").replace("wdl::", "")
    return name


def trace(path: str, top: int) -> str:
    rows = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r.get("Kind", "KERNEL_DISPATCH") != "KERNEL_DISPATCH":
                continue
            wg = int(r["Workgroup_Size_X"]) * int(r["Workgroup_Size_Y"]) * int(r["Workgroup_Size_Z"])
            grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
            dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            rows[(_short(r["Kernel_Name"]), grid // max(wg, 1), wg)].append(dur)
    total = sum(sum(v) for v in rows.values())
    out = ["| kernel | workgroups | threads/WG | calls | avg us | min us | max us | % time |",
           "|---|---|---|---|---|---|---|---|"]
    for (k, nwg, wg), d in sorted(rows.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append("| %s | %d | %d | %d | %.2f | %.2f | %.2f | %.1f |" % (
            k, nwg, wg, len(d), sum(d) / len(d) / 1e3, min(d) / 1e3, max(d) / 1e3,
            100.0 * sum(d) / total))
    return "\n".join(out)


def pmc(path: str, top: int) -> str:
    vals = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            wg = int(r["Workgroup_Size"])
            key = (_short(r["Kernel_Name"]), int(r["Grid_Size"]) // max(wg, 1), r["Counter_Name"])
            vals[key].append(float(r["Counter_Value"]))
    out = ["| kernel | workgroups | counter | dispatches | avg per dispatch | min | max |",
           "|---|---|---|---|---|---|---|"]
    for (k, nwg, c), v in sorted(vals.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append("| %s | %d | %s | %d | %.1f | %.1f | %.1f |" % (
            k, nwg, c, len(v), sum(v) / len(v), min(v), max(v)))
    return "\n".join(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["trace", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args(argv)
    print(trace(a.csv, a.top) if a.kind == "trace" else pmc(a.csv, a.top))
    return 0


if __name__ == "__main__":
    sys.exit(main())
