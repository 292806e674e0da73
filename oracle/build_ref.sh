#!/bin/sh
# Compiles the reference's GSL-free sources on the hot path (Gaunt
# coefficients by Xu's recursion; 2D lattice helpers and point generators)
# straight from /root/reference into oracle/_ref/libqpms_ref.so. The result
# is test infrastructure: it pins the oracle (tests/test_oracle_pins.py) and is
# never linked into the product. The rest of the reference path needs GSL,
# LAPACKE/CBLAS and Amos, which this image lacks, so it is not built.
set -e
REF=${QPMS_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
[ -d "$REF/qpms" ] || { echo "reference sources not present; skipping"; exit 0; }
mkdir -p "$OUT/obj"
DEFS="-DLATTICESUMS32 -DQPMS_VECTORS_NICE_TRANSFORMATIONS -DEWALD_AUTO_CUTOFF"
# lll.c needs cblas.h (absent): l3d_reduceBasis stays unresolved; load with RTLD_LAZY.
for f in gaunt latticegens lattices2d error; do
  gcc -std=gnu11 -fPIC -O2 $DEFS -I"$REF/qpms" -c "$REF/qpms/$f.c" -o "$OUT/obj/$f.o"
done
gcc -shared -o "$OUT/libqpms_ref.so" "$OUT"/obj/*.o -lm
rm -rf "$OUT/obj"
echo "built $OUT/libqpms_ref.so"
