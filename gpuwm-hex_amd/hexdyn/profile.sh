#!/bin/bash
# Kernel-trace and HBM-traffic profile of bench.py (the plain step) with
# rocprofv3, summarised by hexdyn.profile_summary into
# OUTDIR/kernels_<workload>.md and OUTDIR/traffic_<workload>.json (copied into
# profiles/ by hand).
#
#   bash gpuwm-hex_amd/hexdyn/profile.sh [WORKLOAD] [OUTDIR]
#
# Three runs of the same bench command: --kernel-trace --stats (durations),
# --pmc FETCH_SIZE and --pmc WRITE_SIZE (KiB per dispatch; each pass in a
# run of its own).  The counter runs also perform a 256 MiB device copy,
# whose counter values calibrate KiB to bytes.  A first, unprofiled bench
# run tunes the launch tiles and stores them (HEXDYN_CACHE), so the profiled
# runs load the table instead of tuning: large steps per profiled run are
# warmup 3 + timed 8 + profiled 2 = 13.
set -o pipefail
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
WL=${1:-x1.40962-moist}
OUT=${2:-$ROOT/gpurun_out/prof_$WL}
mkdir -p "$OUT"
BENCH="python3 $ROOT/bench.py --workload $WL --steps 8 --warmup 3 --profile-steps 2 --no-cpu-baseline --form plain"
export HEXDYN_CACHE=$OUT/cache
echo "# tune" && timeout -k 10 300 $BENCH > "$OUT/tune.log" 2>&1 || exit $?
cd /tmp && export TMPDIR=/tmp
KT="rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/kt -o kt -- $BENCH"
FE="rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/fetch -o fetch -- $BENCH --calibration-copy-mib 256"
WR="rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/write -o write -- $BENCH --calibration-copy-mib 256"
echo "# kernel trace" && timeout -k 10 300 $KT > "$OUT/kt.log" 2>&1 || exit $?
echo "# FETCH_SIZE" && timeout -s KILL 300 $FE > "$OUT/fetch.log" 2>&1 || exit $?
echo "# WRITE_SIZE" && timeout -s KILL 300 $WR > "$OUT/write.log" 2>&1 || exit $?
find_one() { find "$OUT/$1" -name "$2" | head -n 1; }
cd "$ROOT" && python3 "$ROOT/gpuwm-hex_amd/hexdyn/profile_summary.py" \
  --trace "$(find_one kt 'kt_kernel_stats.csv')" \
  --fetch "$(find_one fetch 'fetch_counter_collection.csv')" \
  --write "$(find_one write 'write_counter_collection.csv')" \
  --steps 13 --workload "$WL" \
  --out-md "$OUT/kernels_$WL.md" --out-json "$OUT/traffic_$WL.json" \
  --command "export HEXDYN_CACHE=$OUT/cache; $BENCH   # tunes and stores the launch tiles" \
  --command "cd /tmp && export TMPDIR=/tmp" --command "$KT" --command "$FE" --command "$WR" \
  --command "python3 gpuwm-hex_amd/hexdyn/profile_summary.py ... --steps 13 (as in gpuwm-hex_amd/hexdyn/profile.sh)"
