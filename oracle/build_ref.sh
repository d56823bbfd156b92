#!/bin/sh
# Builds oracle/_ref/: the reference's GSL-free step headers, compiled in place
# from the reference tree with oracle/ref_harness.cpp, one binary per
# compile-time profile.  Needs the reference tree (this container only) and
# MPICH's headers and library under /opt/conda.  Outputs go to oracle/_ref/.
set -e
REF=${VIRAL_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
[ -d "$REF/NEWEST10" ] || { echo "reference tree not found at $REF; skipping oracle/_ref"; exit 0; }
mkdir -p "$OUT"
MPI_INC=/opt/conda/include
MPI_LIB=/opt/conda/lib
CXX=${CXX:-g++}
# -std=gnu++98: NEWEST10/init.h:11 declares a global 'rank' that clashes with
# std::rank under C++11 and later.  -ffp-contract=off: no FMA contraction.
# --gc-sections drops init() and the other never-called functions that name
# root.h's CheckRoot, the one GSL-side symbol the included headers reference.
# -iquote: only #include "..." looks in the reference tree (its time.h would
# otherwise shadow <time.h>).
FLAGS="-std=gnu++98 -O2 -ffp-contract=off -w -ffunction-sections -fdata-sections -Wl,--gc-sections"
build() {
  name=$1; shift
  $CXX $FLAGS "$@" -iquote "$REF/NEWEST10" -I"$MPI_INC" "$HERE/ref_harness.cpp" -o "$OUT/$name" \
    -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpi -lm
}
build ref_kt -DKT
build ref_kt_gubser -DKT -DGUBSER
build ref_kt_ginit -DKT -DGINIT
build ref_shas -DSHAS
build ref_shas_ginit -DSHAS -DGINIT
build ref_kt_bulktest -DKT -DBULKTEST
echo "built $(ls "$OUT" | tr '\n' ' ')"
