#!/bin/bash
# Builds oracle/_ref/libref_kdtree.so from the reference's src/kdtree2.f90 (compiled where
# it lies, unchanged) and oracle/ref_kdtree_wrap.f90.  Only kdtree2.f90 is buildable here:
# the other modules of the path need generated code (src/gitversion.h), FFTW or NetCDF.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF=/root/reference/src
OUT="$HERE/_ref"
FC=${FC:-flang}
mkdir -p "$OUT"
cd "$OUT"
"$FC" -O2 -fPIC -c "$REF/kdtree2.f90" -o kdtree2.o
"$FC" -O2 -fPIC -c "$HERE/ref_kdtree_wrap.f90" -o ref_kdtree_wrap.o
"$FC" -shared -fPIC kdtree2.o ref_kdtree_wrap.o -o libref_kdtree.so
rm -f kdtree2.o ref_kdtree_wrap.o
echo "built $OUT/libref_kdtree.so"
