#!/bin/bash
# Compiles the reference's benchmark drivers unchanged against this
# repository's headers and links them with the library:
#   tests/cpp/build_reference_drivers.sh OUTDIR [REFERENCE_SRC]
# OUTDIR receives a source tree whose benchmarks/ holds links to the
# reference's rings.{h,cpp} and tensile.{h,cpp}, and whose other headers
# (the includes "../particle.h", "../body.h", ..., "neighbor_search.h",
# "utils.h" of those files) forward to include/mfree; then the executable
# OUTDIR/reference_drivers.  Nothing of the reference is copied.
set -e
OUT=$1
REF=${2:-/root/reference/src}
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
mkdir -p "$OUT/src/benchmarks"
for f in rings.h rings.cpp tensile.h tensile.cpp; do
  ln -sf "$REF/benchmarks/$f" "$OUT/src/benchmarks/$f"
done
for h in particle body correctors cont_mech material derivatives godunov riemann_integration convex_hull \
         boundary interpolation corotational potential; do
  echo "#include \"$ROOT/include/mfree/$h.h\"" > "$OUT/src/$h.h"
done
for h in neighbor_search utils; do
  echo "#include \"$ROOT/include/mfree/benchmarks/$h.h\"" > "$OUT/src/benchmarks/$h.h"
done
CXX=${CXX:-g++}
FLAGS="-std=c++14 -O1 -I$ROOT/include"
$CXX $FLAGS -c "$OUT/src/benchmarks/rings.cpp" -o "$OUT/rings.o"
$CXX $FLAGS -c "$OUT/src/benchmarks/tensile.cpp" -o "$OUT/tensile.o"
$CXX $FLAGS -I"$OUT/src/benchmarks" -c "$ROOT/tests/cpp/reference_drivers_main.cpp" -o "$OUT/main.o"
$CXX -o "$OUT/reference_drivers" "$OUT/main.o" "$OUT/rings.o" "$OUT/tensile.o" -L"$ROOT/mfree_iwf_amd" \
  -lmfree_iwf_amd -Wl,-rpath,"$ROOT/mfree_iwf_amd"
