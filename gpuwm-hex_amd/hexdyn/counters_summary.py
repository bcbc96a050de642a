"""Summarise rocprofv3 --pmc runs of ``bench.py`` into a per-kernel table of
memory-pipeline counters (mean per launch; per-CU figures divide the block
totals by the CU count).

    python -m hexdyn.counters_summary --csv DIR/l2_counter_collection.csv --csv ... \\
        --kernels "k_a|k_b" --workload NAME --out-md profiles/counters_NAME.md [--command "..."]...
"""

from __future__ import annotations

import argparse
import collections
import csv
from pathlib import Path

from profile_summary import short_name  # noqa: E402  (run as a script beside it)

CUS = 256


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--csv", action="append", required=True, type=Path)
    ap.add_argument("--kernels", required=True)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--out-md", required=True, type=Path)
    ap.add_argument("--command", action="append", default=[])
    a = ap.parse_args()
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in a.csv:
        for row in csv.DictReader(open(path)):
            vals[short_name(row["Kernel_Name"])][row["Counter_Name"]].append(float(row["Counter_Value"]))
    kernels = [k for k in a.kernels.split("|") if k in vals]

    def m(k, c):
        v = vals[k].get(c)
        return sum(v) / len(v) if v else float("nan")

    lines = [f"# Memory-pipeline counters of the gather kernels, {a.workload} (plain step, strict numerics)", "",
             "Commands (repository root on the GPU box, TMPDIR=/tmp), one pass each:", ""]
    lines += ["    " + c for c in a.command]
    lines += ["", f"Per launch (mean over launches; per-CU figures divide block counters by {CUS} CUs).", "",
              "| kernel | waves | vmem rd / wave | vmem wr / wave | L2 hit % | L1 accesses / vmem instr | "
              "TA busy cyc/CU | TD busy cyc/CU | TCP pending-stall cyc/CU | wave-cycles waiting / wave-cycles |",
              "|---|---|---|---|---|---|---|---|---|---|"]
    for k in kernels:
        waves = m(k, "SQ_WAVES")
        rd, wr = m(k, "SQ_INSTS_VMEM_RD"), m(k, "SQ_INSTS_VMEM_WR")
        hit, miss = m(k, "TCC_HIT_sum"), m(k, "TCC_MISS_sum")
        acc = m(k, "TCP_TOTAL_CACHE_ACCESSES_sum")
        wait, wcyc = m(k, "SQ_WAIT_INST_ANY"), m(k, "SQ_WAVE_CYCLES")
        lines.append(f"| {k} | {waves:.0f} | {rd / waves:.1f} | {wr / waves:.1f} | {100 * hit / (hit + miss):.0f} | "
                     f"{acc / (rd + wr):.2f} | {m(k, 'TA_TA_BUSY_sum') / CUS:.0f} | {m(k, 'TD_TD_BUSY_sum') / CUS:.0f} | "
                     f"{m(k, 'TCP_PENDING_STALL_CYCLES_sum') / CUS:.0f} | {wait / wcyc:.2f} |")
    a.out_md.write_text("\n".join(lines) + "\n")
    print("\n".join(lines[-len(kernels) - 2:]))


if __name__ == "__main__":
    main()
