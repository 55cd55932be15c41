#!/bin/bash
# Builds the reference's own CPU code (accurate_FSM.cpp, classic.cpp, the
# acoustic modelling functions of utils/wave/acoustic.cpp and the utilities
# they need) from the reference checkout, with the driver
# oracle/ref_driver.cpp, into oracle/_ref/libstref.so.  Test infrastructure:
# it pins the restatement in oracle/oracle.c and can time the reference CPU
# path.  Output goes only to oracle/_ref/; nothing is written under REF.
# The reference's CUDA block FIM and least_squares.cpp (which includes the
# CUDA header chain) are not buildable here and are not attempted.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${ST_REFERENCE:-/root/reference}"
if [ ! -d "$REF/src" ]; then
    echo "build_ref.sh: reference checkout not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
SRC="$REF/src"
g++ -std=c++17 -O2 -ffp-contract=off -fPIC -shared -Wno-unknown-pragmas -w \
    -I"$SRC" \
    "$SRC/eikonal/accurate_FSM/accurate_FSM.cpp" \
    "$SRC/eikonal/classic/classic.cpp" \
    "$SRC/utils/file_manager/file_manager.cpp" \
    "$SRC/utils/interpolation/trilinear.cpp" \
    "$SRC/utils/smoothing/gaussian.cpp" \
    "$SRC/utils/wave/acoustic.cpp" \
    "$SRC/geometry/regular/regular.cpp" \
    "$SRC/geometry/circular/circular.cpp" \
    "$HERE/ref_driver.cpp" \
    -o "$HERE/_ref/libstref.so.tmp" -lpthread
mv "$HERE/_ref/libstref.so.tmp" "$HERE/_ref/libstref.so"
echo "built $HERE/_ref/libstref.so"
