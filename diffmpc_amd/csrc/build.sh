#!/usr/bin/env bash
# Build libdmpc.so for gfx950 in-tree (next to the Python package).  The
# translation units compile in parallel, then link with --no-undefined.
# Packed-FP32 instructions are disabled on purpose: hipcc -O3 would otherwise
# emit v_pk_* FP32 forms by itself.  Loops start on 64-byte boundaries: the
# PCG loops' speed moved by about 1 % with their start address otherwise.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
out="${DMPC_LIB_OUT:-${here}/../libdmpc.so}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
obj="$(mktemp -d)"
trap 'rm -rf "${obj}"' EXIT
flags=(-O3 -std=c++17 --offload-arch=gfx950 -fPIC -falign-loops=64
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -mcode-object-version=5 -Wall -Wno-unused-function)
# custom build: DMPC_CUSTOM_MODEL_FILE names a header defining dmpc::CustomModel
if [ -n "${DMPC_CUSTOM_MODEL_FILE:-}" ]; then
  flags+=("-DDMPC_CUSTOM_MODEL_FILE=\"${DMPC_CUSTOM_MODEL_FILE}\"")
fi
pids=()
for src in dmpc_fused dmpc_stream dmpc_plain dmpc_dyngrad dmpc_simulate dmpc_api; do
  "${HIPCC}" "${flags[@]}" -c "${here}/${src}.hip" -o "${obj}/${src}.o" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "${p}"; done
"${HIPCC}" --offload-arch=gfx950 -fPIC -shared "${obj}"/*.o \
  -Wl,--no-undefined -L"${ROCM_PATH:-/opt/rocm}/lib" -lamdhip64 \
  -o "${out}.tmp"
mv "${out}.tmp" "${out}"
echo "built ${out}"
