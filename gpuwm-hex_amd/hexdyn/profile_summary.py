"""Summarise rocprofv3 output of ``bench.py`` into a per-kernel table.

    python -m hexdyn.profile_summary --trace DIR/kt_kernel_stats.csv \\
        --fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \\
        --steps N --workload NAME --out-md profiles/X.md --out-json profiles/X.json \\
        [--counters DIR/sq_counter_collection.csv] [--command "..."]...

The kernel-trace statistics give calls and durations (divided by the number
of large steps the traced run executed).  FETCH_SIZE and WRITE_SIZE are
converted to bytes with the factor that makes the 256 MiB device copy the
profiled run performs read and write exactly 256 MiB (``bench.py
--calibration-copy-mib 256``).  Optional SQ counters (instruction counts and
busy cycles) are averaged per dispatch.  The JSON maps each kernel to its
average HBM bytes per launch; ``bench.py`` reads it for ``roofline.traffic``.
"""

from __future__ import annotations

import argparse
import collections
import csv
import json
import re
from pathlib import Path

COPY_KIB = 256 * 1024


def short_name(name: str) -> str:
    m = re.search(r"::(k_\w+)(?:<\d+>)?\(", name)
    if m:
        return m.group(1)
    m = re.search(r"(k_\w+)", name)
    return m.group(1) if m else name


def per_dispatch(path: Path, counter: str):
    agg = collections.defaultdict(list)
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] == counter:
            agg[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    return agg


def calibration(agg, minimum_kib: float) -> float:
    vals = [v for n, vs in agg.items() if "copyBuffer" in n for v in vs if v > minimum_kib]
    if not vals:
        raise SystemExit("no calibration copy found in the counter file")
    return COPY_KIB / (sum(vals) / len(vals))


def mean_by_kernel(agg):
    out = collections.defaultdict(list)
    for name, vals in agg.items():
        out[short_name(name)].extend(vals)
    return {k: sum(v) / len(v) for k, v in out.items()}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True, type=Path)
    ap.add_argument("--fetch", required=True, type=Path)
    ap.add_argument("--write", required=True, type=Path)
    ap.add_argument("--counters", type=Path)
    ap.add_argument("--steps", required=True, type=float)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--out-md", required=True, type=Path)
    ap.add_argument("--out-json", required=True, type=Path)
    ap.add_argument("--command", action="append", default=[])
    a = ap.parse_args()

    fetch = per_dispatch(a.fetch, "FETCH_SIZE")
    write = per_dispatch(a.write, "WRITE_SIZE")
    ff = calibration(fetch, COPY_KIB * 0.25)
    fw = calibration(write, COPY_KIB * 0.5)
    rd, wr = mean_by_kernel(fetch), mean_by_kernel(write)
    sq = {}
    if a.counters:
        by = collections.defaultdict(lambda: collections.defaultdict(list))
        for row in csv.DictReader(open(a.counters)):
            by[short_name(row["Kernel_Name"])][row["Counter_Name"]].append(float(row["Counter_Value"]))
        sq = {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in by.items()}

    rows = [r for r in csv.DictReader(open(a.trace)) if "hexdyn" in r["Name"]]
    merged = collections.defaultdict(lambda: [0.0, 0.0])  # calls, total ns
    for r in rows:
        m = merged[short_name(r["Name"])]
        m[0] += float(r["Calls"])
        m[1] += float(r["TotalDurationNs"])
    total = sum(v[1] for v in merged.values())
    lines = [f"# Per-kernel time and HBM traffic, {a.workload}", "",
             "Commands (repository root on the GPU box, TMPDIR=/tmp):", ""]
    lines += [f"    {c}" for c in a.command]
    lines += ["", f"Large steps in the kernel-trace run: {a.steps:g}.",
              f"Counter calibration: FETCH_SIZE x {ff:.3f} and WRITE_SIZE x {fw:.3f} give bytes "
              "(a 256 MiB device copy reads and writes 256 MiB).", ""]
    head = "| kernel | calls/step | avg us | ms/step | % | HBM read MB/launch | HBM write MB/launch | TB/s |"
    if sq:
        head += " VALU/wave | SALU+SMEM/wave |"
    lines += [head, "|" + "---|" * (head.count("|") - 1)]
    traffic = {}
    for k, (calls, tns) in sorted(merged.items(), key=lambda kv: -kv[1][1]):
        avg_us = tns / calls / 1e3
        r_mb = rd.get(k, float("nan")) * ff * 1024 / 1e6
        w_mb = wr.get(k, float("nan")) * fw * 1024 / 1e6
        traffic[k] = (r_mb + w_mb) * 1e6
        line = (f"| {k} | {calls / a.steps:g} | {avg_us:.1f} | {tns / a.steps / 1e6:.3f} | {100 * tns / total:.1f} | "
                f"{r_mb:.1f} | {w_mb:.1f} | {(r_mb + w_mb) / avg_us:.2f} |")
        if sq:
            d = sq.get(k, {})
            waves = d.get("SQ_WAVES", float("nan"))
            line += (f" {d.get('SQ_INSTS_VALU', float('nan')) / waves:.0f} |"
                     f" {(d.get('SQ_INSTS_SALU', float('nan')) + d.get('SQ_INSTS_SMEM', 0.0)) / waves:.0f} |")
        lines.append(line)
    lines += ["", f"Sum of hexdyn kernel time per large step: {total / a.steps / 1e6:.3f} ms."]
    a.out_md.write_text("\n".join(lines) + "\n")
    a.out_json.write_text(json.dumps({"workload": a.workload, "hbm_bytes_per_launch": traffic}, indent=1) + "\n")
    print("\n".join(lines[:40]))


if __name__ == "__main__":
    main()
