#!/usr/bin/env bash
# AddressSanitizer build of the C ABI (host code instrumented, device code as
# in build.sh) and of the argument-check driver check_abi.c; runs the driver.
# CPU only: the driver makes no call that reaches the GPU.
# Usage: build_check.sh [out_dir]   (default: diffmpc_amd/csrc/asan/out)
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
csrc="${here}/.."
out="${1:-${here}/out}"
ROCM="${ROCM_PATH:-/opt/rocm}"
mkdir -p "${out}"
obj="$(mktemp -d)"
trap 'rm -rf "${obj}"' EXIT
flags=(-O2 -std=c++17 --offload-arch=gfx950 -fPIC
       -Xclang -target-feature -Xclang -packed-fp32-ops -mcode-object-version=5
       -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer -Xarch_host -g)
pids=()
for src in dmpc_fused dmpc_stream dmpc_plain dmpc_dyngrad dmpc_simulate dmpc_api; do
  "${ROCM}/bin/hipcc" "${flags[@]}" -c "${csrc}/${src}.hip" -o "${obj}/${src}.o" 2>/dev/null &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "${p}"; done
"${ROCM}/llvm/bin/clang++" -shared -fsanitize=address "${obj}"/*.o -L"${ROCM}/lib" -lamdhip64 \
  -Wl,-rpath,"${ROCM}/lib" -o "${out}/libdmpc_asan.so"
"${ROCM}/llvm/bin/clang" -fsanitize=address -fno-omit-frame-pointer -g -I"${csrc}/../../include" \
  "${here}/check_abi.c" -L"${out}" -ldmpc_asan -Wl,-rpath,"${out}" -o "${out}/check_abi"
# the HIP runtime's own allocations at exit are not this library's
ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1" "${out}/check_abi"
