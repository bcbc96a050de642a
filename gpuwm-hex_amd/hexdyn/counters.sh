#!/bin/bash
# Memory-pipeline counters of the hot gather kernels of bench.py (the plain
# step), four rocprofv3 --pmc passes (each within the per-block counter
# limits, each in a run of its own), summarised by hexdyn.counters_summary
# into OUTDIR/counters_<workload>.md (copied into profiles/ by hand).
#
#   bash gpuwm-hex_amd/hexdyn/counters.sh [WORKLOAD] [OUTDIR]
#
# A first, unprofiled bench run tunes and stores the launch tiles
# (HEXDYN_CACHE) so the counter runs load them instead of tuning.
set -o pipefail
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
WL=${1:-x1.40962-moist}
OUT=${2:-$ROOT/gpurun_out/pmc_$WL}
mkdir -p "$OUT"
BENCH="python3 $ROOT/bench.py --workload $WL --steps 2 --warmup 0 --profile-steps 1 --no-cpu-baseline --form plain"
# the edge-flux pass under both names: its tiled form runs unless a launch
# table selects the gather form
KERNELS="k_diag_edge_flux_t|k_diag_edge_flux|k_tend_mom_ac|k_rec_cell|k_ac_cell|k_tend_cell|k_rec_edge|k_sc_edge|k_diag_cell"
export HEXDYN_CACHE=$OUT/cache
echo "# tune" && timeout -k 10 300 $BENCH > "$OUT/tune.log" 2>&1 || exit $?
cd /tmp && export TMPDIR=/tmp
P1="rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum TCC_REQ_sum --kernel-include-regex $KERNELS --output-format csv -d $OUT/l2 -o l2 -- $BENCH"
P2="rocprofv3 --pmc TCP_TCC_READ_REQ_sum TCP_TOTAL_CACHE_ACCESSES_sum TCP_PENDING_STALL_CYCLES_sum TCP_TCR_TCP_STALL_CYCLES_sum --kernel-include-regex $KERNELS --output-format csv -d $OUT/l1 -o l1 -- $BENCH"
P3="rocprofv3 --pmc SQ_INST_LEVEL_VMEM SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_WAVE_CYCLES SQ_WAIT_INST_ANY SQ_BUSY_CYCLES SQ_WAVES SQ_INSTS_VALU --kernel-include-regex $KERNELS --output-format csv -d $OUT/sq -o sq -- $BENCH"
P4="rocprofv3 --pmc TA_TA_BUSY_sum TA_BUFFER_READ_WAVEFRONTS_sum TD_TD_BUSY_sum TD_SPI_STALL_sum --kernel-include-regex $KERNELS --output-format csv -d $OUT/ta -o ta -- $BENCH"
echo "# L2" && timeout -s KILL 200 $P1 > "$OUT/l2.log" 2>&1 || exit $?
echo "# L1" && timeout -s KILL 200 $P2 > "$OUT/l1.log" 2>&1 || exit $?
echo "# SQ" && timeout -s KILL 200 $P3 > "$OUT/sq.log" 2>&1 || exit $?
echo "# TA/TD" && timeout -s KILL 200 $P4 > "$OUT/ta.log" 2>&1 || exit $?
find_one() { find "$OUT/$1" -name "$2" | head -n 1; }
cd "$ROOT" && python3 "$ROOT/gpuwm-hex_amd/hexdyn/counters_summary.py" \
  --csv "$(find_one l2 'l2_counter_collection.csv')" --csv "$(find_one l1 'l1_counter_collection.csv')" \
  --csv "$(find_one sq 'sq_counter_collection.csv')" --csv "$(find_one ta 'ta_counter_collection.csv')" \
  --kernels "$KERNELS" --workload "$WL" --out-md "$OUT/counters_$WL.md" \
  --command "export HEXDYN_CACHE=$OUT/cache; $BENCH   # tunes and stores the launch tiles" \
  --command "cd /tmp && export TMPDIR=/tmp" --command "$P1" --command "$P2" --command "$P3" --command "$P4"
