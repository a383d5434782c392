#!/usr/bin/env bash
# Build libdmpc.so for gfx950 in-tree (next to the Python package).
# Packed-FP32 instructions are disabled on purpose: hipcc -O3 would otherwise
# emit v_pk_* FP32 forms by itself.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
out="${here}/../libdmpc.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"${HIPCC}" -O3 -std=c++17 --offload-arch=gfx950 -fPIC -shared \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -mcode-object-version=5 \
  -Wall -Wno-unused-function \
  "${here}/dmpc_fused.hip" "${here}/dmpc_stream.hip" "${here}/dmpc_plain.hip" "${here}/dmpc_api.hip" \
  -Wl,--no-undefined -L"${ROCM_PATH:-/opt/rocm}/lib" -lamdhip64 \
  -o "${out}.tmp"
mv "${out}.tmp" "${out}"
echo "built ${out}"
