#!/bin/bash
# Builds libjvmc_hip.so for gfx950 in-tree (next to this script's package).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${JV_OUT:-$here/../libjvmc_hip.so}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
flags=(${JV_EXTRA_FLAGS:-} -O3 -std=c++17 -fPIC -shared --offload-arch=gfx950 -mcode-object-version=5
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -Wno-unused-result -Wno-unused-parameter)
srcs=("$here/core.hip" "$here/nets.hip" "$here/eloc.hip" "$here/sampler.hip" "$here/linalg.hip" "$here/tridiag.hip" "$here/sym.hip" "$here/dc.hip" "$here/chol.hip" "$here/gram_mixed.hip" "$here/batch.hip")
"$HIPCC" "${flags[@]}" "${srcs[@]}" -o "$out.tmp" -L/opt/rocm/lib -lrocsolver -lrocblas
mv "$out.tmp" "$out"
echo "built $out"
