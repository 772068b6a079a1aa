#!/usr/bin/env bash
# Builds libmr3hip.so for gfx950 in place (mr3smp_amd/libmr3hip.so).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/../libmr3hip.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
# FP64 throughout, no contraction (Sturm counts must round like the reference),
# no packed-FP32 forms.
FLAGS="--offload-arch=gfx950 -O3 -fPIC -std=c++17 -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result"
"$HIPCC" $FLAGS -shared -o "$OUT.tmp" "$HERE/mr3_driver.hip" "$HERE/mr3_dense.hip" \
  "$HERE/mr3_api.cpp" "$HERE/mr3_serve.cpp" -lrocsolver -lrocblas -lpthread
mv -f "$OUT.tmp" "$OUT"
echo "built $OUT"
