#!/bin/bash
# Compiles the self-contained part of the reference's hot path from its own
# sources (no stand-ins): the 13-point derivative tables and Deriv3D_p/_pc
# (init_deriv_parallel.f90, newder.f, redef.f, mates.f) and the LJ transform
# with its sine integral (fforma.f90, s13adf.f90).  Output: oracle/_ref/libref.so
# (git-ignored, used only by the tests as a second checker of the oracle).
# poten/energy/steppc/steprk need FFTW3 and module state and are not built.
set -euo pipefail
REF=${EFFPOT_REFERENCE:-/root/reference}/src
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/../_ref
FC=/opt/rocm/llvm/bin/flang
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"
for f in init_deriv_parallel.f90 newder.f redef.f mates.f fforma.f90 s13adf.f90; do
  $FC -O2 -fPIC -fopenmp -c "$REF/$f" -o "${f%.*}.o"
done
$FC -shared -fopenmp -o "$OUT/libref.so.tmp" ./*.o
mv "$OUT/libref.so.tmp" "$OUT/libref.so"
echo "built $OUT/libref.so"
