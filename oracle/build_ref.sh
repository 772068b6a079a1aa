#!/usr/bin/env bash
# Builds the reference mr3smp library (C + Fortran 77 sources, unmodified, read in place
# under /root/reference) into oracle/_ref/libmrrr_ref.so.  Used only as a checker for the
# oracle restatement and as the CPU baseline; never linked into the product.
# Toolchain: gcc for the C files, ROCm amdflang for the vendored LAPACK/BLAS files.
set -euo pipefail
REF=${MR3_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/bin/amdflang}
if [ ! -d "$REF/SRC" ]; then
  echo "reference sources not found at $REF; skipping _ref build" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
# Same preprocessor settings as the reference Makefile/make.inc: -DNDEBUG, vendored LAPACK
# (-DNOLAPACK), spinlocks on, complex drivers off.
CFLAGS="-O3 -fPIC -pthread -DNDEBUG -DNOLAPACK -I$REF/INCLUDE"
FFLAGS="-O2 -fPIC"
objs=()
for f in "$REF"/SRC/*.c; do
  b=$(basename "$f" .c)
  [ "$b" = "dense" ] && continue   # dense drivers need an external LAPACK; not on this path
  gcc $CFLAGS -c "$f" -o "$OUT/obj/$b.o"
  objs+=("$OUT/obj/$b.o")
done
for f in "$REF"/SRC/LAPACK/*.f "$REF"/SRC/BLAS/*.f; do
  b=$(basename "$f" .f)
  # SRC/BLAS/odswap.f duplicates SRC/LAPACK/odswap.f; a static archive tolerates that, a .so does not
  [ "$f" = "$REF/SRC/BLAS/odswap.f" ] && continue
  "$FC" $FFLAGS -c "$f" -o "$OUT/obj/f_$b.o"
  objs+=("$OUT/obj/f_$b.o")
done
"$FC" -shared -o "$OUT/libmrrr_ref.so" "${objs[@]}" -lpthread -lm
echo "built $OUT/libmrrr_ref.so"
