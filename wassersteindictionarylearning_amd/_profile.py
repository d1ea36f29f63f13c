"""Summaries of rocprofv3 CSV output: kernel durations and PMC counters per
kernel and grid size (rocprofv3's own --stats table groups by name only, which
mixes launches of different shapes -- e.g. the batched run's per-stream
histogram groups and a stage launched over all histograms).

python -m wassersteindictionarylearning_amd._profile trace <..._kernel_trace.csv> [--top N]
python -m wassersteindictionarylearning_amd._profile pmc <..._counter_collection.csv> [--top N]
python -m wassersteindictionarylearning_amd._profile issue <..._counter_collection.csv> [--top N]

Prints Markdown tables: calls, average / min / max duration in microseconds
(trace), the average counter value per dispatch in the counter's unit
(pmc; FETCH_SIZE and WRITE_SIZE are KiB), or issue rates (issue, from a run
collecting SQ_INSTS_VALU, SQ_WAVES, SQ_LDS_IDX_ACTIVE, SQ_LDS_BANK_CONFLICT and
GRBM_GUI_ACTIVE): VALU issue = 4 SQ_INSTS_VALU / (cycles x SIMDs) -- a wave64
VALU instruction holds a 16-lane SIMD for 4 cycles, quarter-rate ones longer, so
it is a lower bound -- LDS busy = SQ_LDS_IDX_ACTIVE / (cycles x CUs), and the
bank-conflict share of the LDS cycles.  GRBM_GUI_ACTIVE is summed over the 8
XCDs (cycles = its value / 8); MI355X: 256 CUs, 1024 SIMDs.
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


def issue(path: str, top: int, xcds: int = 8, cus: int = 256, simds: int = 1024) -> str:
    per = defaultdict(dict)  # dispatch -> counter -> value
    shape = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            wg = int(r["Workgroup_Size"])
            per[r["Dispatch_Id"]][r["Counter_Name"]] = float(r["Counter_Value"])
            shape[r["Dispatch_Id"]] = (_short(r["Kernel_Name"]), int(r["Grid_Size"]) // max(wg, 1), wg)
    agg = defaultdict(lambda: defaultdict(float))
    n = defaultdict(int)
    for d, v in per.items():
        n[shape[d]] += 1
        for c, x in v.items():
            agg[shape[d]][c] += x
    out = ["| kernel | workgroups | threads/WG | dispatches | VALU issue | LDS busy | "
           "bank-conflict share of LDS cycles | VALU instructions per wave |",
           "|---|---|---|---|---|---|---|---|"]
    for key, v in sorted(agg.items(), key=lambda kv: -kv[1]["GRBM_GUI_ACTIVE"])[:top]:
        cyc = v["GRBM_GUI_ACTIVE"] / xcds
        if cyc <= 0:
            continue
        out.append("| %s | %d | %d | %d | %.2f | %.2f | %.3f | %.0f |" % (
            key[0], key[1], key[2], n[key], 4 * v["SQ_INSTS_VALU"] / (cyc * simds),
            v["SQ_LDS_IDX_ACTIVE"] / (cyc * cus),
            v["SQ_LDS_BANK_CONFLICT"] / max(1.0, v["SQ_LDS_IDX_ACTIVE"]),
            v["SQ_INSTS_VALU"] / max(1.0, v["SQ_WAVES"])))
    return "\n".join(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["trace", "pmc", "issue"])
    ap.add_argument("csv")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args(argv)
    fn = {"trace": trace, "pmc": pmc, "issue": issue}[a.kind]
    print(fn(a.csv, a.top))
    return 0


if __name__ == "__main__":
    sys.exit(main())
