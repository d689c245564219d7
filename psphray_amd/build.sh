#!/bin/bash
# Builds psphray_amd/libpsphray_amd.so for gfx950 (in-tree) and the CPU
# oracle (oracle/liboracle.so).  Usage: psphray_amd/build.sh [-j N]
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(dirname "$HERE")"
OUT="$HERE/_build"
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
# -ffp-contract=off: the traversal reproduces separately rounded float
#   products of the reference geometry; -packed-fp32-ops off: packed FP32
#   instructions are left out of every kernel (the compiler otherwise forms
#   them from scalar or float2 code)
HIPFLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -ffp-contract=off -munsafe-fp-atomics
          -Xclang -target-feature -Xclang -packed-fp32-ops
          -I"$ROOT/include" -I"$HERE/csrc" -Wno-unused-result)
# PSR_FORCE=1 recompiles every object; otherwise only stale ones
FORCE="${PSR_FORCE:-0}"
pids=()
for src in "$HERE"/csrc/*.hip; do
  obj="$OUT/$(basename "${src%.hip}").o"
  stale=0
  for dep in "$src" "$HERE"/csrc/*.h "$ROOT/include/psphray_amd.h"; do
    [[ "$dep" -nt "$obj" ]] && stale=1
  done
  # the deposit: no NaN reaches its arithmetic, so fmin/fmax take their
  # operands as they are (without NaN-quieting instructions on the walk's chain)
  extra=()
  [[ "$(basename "$src")" == psr_deposit.hip ]] && extra=(-fno-honor-nans)
  if [[ "$FORCE" == 1 || ! -f "$obj" || "$stale" == 1 ]]; then
    "$HIPCC" "${HIPFLAGS[@]}" "${extra[@]}" -c "$src" -o "$obj" &
    pids+=($!)
  fi
done
gcc -std=gnu99 -O2 -fPIC -ffp-contract=off -I"$ROOT/include" -c "$HERE/host/psr_host.c" -o "$OUT/psr_host.o" &
pids+=($!)
for p in "${pids[@]}"; do wait "$p" || { echo "build.sh: a compile step failed" >&2; exit 1; }; done
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -o "$HERE/libpsphray_amd.so" "$OUT"/*.o -lm
# a plain C caller of the ABI (tests/test_c_driver.py runs it)
mkdir -p "$ROOT/tests/_bin"
gcc -std=c99 -O2 -Wall -I"$ROOT/include" "$ROOT/tests/c_driver.c" -L"$HERE" -lpsphray_amd \
    -Wl,-rpath,'$ORIGIN/../../psphray_amd' -o "$ROOT/tests/_bin/c_driver"
make -s -C "$ROOT/oracle" all
