#!/bin/bash
# Compiles the reference's benchmark drivers unchanged against this
# repository's headers and links them with the library:
#   tests/cpp/build_reference_drivers.sh OUTDIR [REFERENCE_SRC]
# OUTDIR receives a source tree whose benchmarks/ holds links to the
# reference's rings.{h,cpp}, tensile.{h,cpp} and rotation.h, whose top level
# links the reference's main program (mfree_iwf.cpp), and whose other headers
# (the includes "../particle.h", "../body.h", ..., "neighbor_search.h",
# "utils.h", "contact.h", "vtk_writer.h" of those files) forward to
# include/mfree; then the executables OUTDIR/reference_drivers (this
# repository's probe main over the reference's drivers) and
# OUTDIR/reference_main (the reference's own main, unchanged; it steps
# forever and is only linked here).  Nothing of the reference is copied.
set -e
OUT=$1
REF=${2:-/root/reference/src}
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
mkdir -p "$OUT/src/benchmarks"
for f in rings.h rings.cpp tensile.h tensile.cpp rotation.h; do
  ln -sf "$REF/benchmarks/$f" "$OUT/src/benchmarks/$f"
done
ln -sf "$REF/mfree_iwf.cpp" "$OUT/src/mfree_iwf.cpp"
for h in particle body correctors cont_mech material derivatives godunov riemann_integration convex_hull \
         boundary interpolation corotational potential contact vtk_writer; do
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
$CXX $FLAGS -c "$OUT/src/mfree_iwf.cpp" -o "$OUT/mfree_iwf.o"
$CXX -o "$OUT/reference_main" "$OUT/mfree_iwf.o" "$OUT/rings.o" "$OUT/tensile.o" -L"$ROOT/mfree_iwf_amd" \
  -lmfree_iwf_amd -Wl,-rpath,"$ROOT/mfree_iwf_amd"
