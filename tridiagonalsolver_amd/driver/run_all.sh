#!/bin/bash
# Algorithm x size sweep of the reference (Tridiagonal_CUDA/run_all.sh:12-48) against
# libtridiag_amd: every run solves MAXSIZE unknowns (batch = MAXSIZE / n), REP
# repetitions, and appends one CSV line per run (columns of test.cpp:121-133).
#
#   REP=200 ./run_all.sh            # the reference settings
#   SIZES="512 4096" REP=5 OUT=/tmp/res ./run_all.sh
set -o pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
DRIVER=${DRIVER:-$HERE/../lib/tridiag_test}
SUFFIX=$(date +%m-%d-%H%M)
MAXSIZE=${MAXSIZE:-134217728}
REP=${REP:-200}
OUT=${OUT:-.}
read -r -a sizes <<< "${SIZES:-64 128 192 256 320 384 448 512 1024 1536 2048 2560 3072 3584 4096 8192 12288 16384 20480 24576 28672 32768 65536}"
read -r -a algos_single <<< "${ALGOS_SINGLE:-0 1 3 4 5 6 9 14}"
read -r -a algos_double <<< "${ALGOS_DOUBLE:-0 1 3 4 9}"

status=0
RES=$OUT/result_single_$SUFFIX.csv
for algo in "${algos_single[@]}"; do
    for n in "${sizes[@]}"; do
        "$DRIVER" 0 "$algo" "$n" $((MAXSIZE / n)) "$REP" "$RES" || status=1
    done
done
RES=$OUT/result_double_$SUFFIX.csv
for algo in "${algos_double[@]}"; do
    for n in "${sizes[@]}"; do
        "$DRIVER" 1 "$algo" "$n" $((MAXSIZE / n)) "$REP" "$RES" || status=1
    done
done
exit $status
