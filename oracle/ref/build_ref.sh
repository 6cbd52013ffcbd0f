#!/usr/bin/env bash
# build_ref.sh -- compile the reference solver's own hot-path routines into
# oracle/_ref/libref_<tag>.so (TEST INFRASTRUCTURE ONLY).
#
#   build_ref.sh X_OUT Y_OUT Z_OUT NP NX_INLET NX_OUTLET
#
# The reference fixes its grid at compile time in Resolution.h and its rank
# count in MPI_prm.h.  This script writes resized copies of those two
# configuration headers into a temporary directory outside the repository,
# then compiles the reference .cpp files where they lie under the reference
# tree with `-I<tmp> -I-`, so that the quoted includes resolve to the resized
# headers.  Nothing from the reference is written into the repository; only
# the linked library lands in oracle/_ref/.  Flags follow the reference's
# GNU Makefile (Makefile:9: mpicxx -O3, no OpenMP); with CGLI_REF_OPENMP=1 the
# routines are compiled with -fopenmp as in the reference's tuned build
# (Makefile_K:12), into libref_<tag>_omp.so, for timing only.
set -euo pipefail
REF=${CGLI_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref
X=$1; Y=$2; Z=$3; NP=$4; NIN=$5; NOUT=$6
TAG="${X}x${Y}x${Z}_np${NP}_in${NIN}_out${NOUT}"
OMPF=""
if [ "${CGLI_REF_OPENMP:-0}" = 1 ]; then OMPF="-fopenmp"; TAG="${TAG}_omp"; fi
if [ ! -f "$REF/Resolution.h" ]; then echo "reference tree not found at $REF" >&2; exit 3; fi
MPIINC=/opt/conda/include; MPILIB=/opt/conda/lib
mkdir -p "$OUT"
TMP=$(mktemp -d /tmp/cgli_ref_XXXXXX)
trap 'rm -rf "$TMP"' EXIT
sed -e "s/^#define X_out [0-9]*/#define X_out $X/" \
    -e "s/^#define Y_out [0-9]*/#define Y_out $Y/" \
    -e "s/^#define Z_out [0-9]*/#define Z_out $Z/" \
    -e "s/^#define X_m [0-9]*/#define X_m $((X+4))/" \
    -e "s/^#define Y_m [0-9]*/#define Y_m $((Y+4))/" \
    -e "s/^#define Z_m [0-9]*/#define Z_m $((Z+4))/" \
    -e "s/^#define nx_inlet *[0-9]*/#define nx_inlet $NIN/" \
    -e "s/^#define nx_outlet *[0-9]*/#define nx_outlet $NOUT/" \
    "$REF/Resolution.h" > "$TMP/Resolution.h"
sed -e "s/^#define np [0-9]*/#define np $NP/" "$REF/MPI_prm.h" > "$TMP/MPI_prm.h"
grep -q "#define X_out $X\b" "$TMP/Resolution.h" && grep -q "#define np $NP\b" "$TMP/MPI_prm.h" || { echo "header resize failed" >&2; exit 4; }
INC="-I$TMP -I- -I$REF -I$MPIINC"
TUS="Flux_X Flux_Y Flux_Z Viscous_terms Dual_time_stepping Hyperplane Jacobian Jacobian_u Jacobian_v Jacobian_w Y_boundary_condition Z_boundary_condition Filter Output Initial_condition"
for t in $TUS; do
  g++ -O3 -fPIC -w $OMPF $INC -c "$REF/$t.cpp" -o "$TMP/$t.o"
done
g++ -O3 -fPIC -w $INC -c "$HERE/ref_harness.cpp" -o "$TMP/ref_harness.o"
g++ -shared $OMPF -o "$OUT/libref_${TAG}.so" "$TMP"/*.o -L$MPILIB -Wl,-rpath,$MPILIB -lmpi
echo "$OUT/libref_${TAG}.so"
