#!/bin/bash
# Builds qpanda_amd/libqsv.so (HIP kernels + C ABI) for gfx950.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
cd "$HERE"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops
       -Wall -Wno-unused-function -I"$HERE/../include")
mkdir -p _build
OBJS=()
PIDS=()
for src in csrc/*.hip csrc/*.cpp; do
  obj="_build/$(basename "$src").o"
  OBJS+=("$obj")
  if [[ "$src" == *.hip ]]; then
    "$HIPCC" "${FLAGS[@]}" -c "$src" -o "$obj" 2> "$obj.log" &
  else
    "$HIPCC" "${FLAGS[@]}" -x hip -c "$src" -o "$obj" 2> "$obj.log" &
  fi
  PIDS+=($!)
done
fail=0
for i in "${!PIDS[@]}"; do
  if ! wait "${PIDS[$i]}"; then
    fail=1
  fi
  grep -v -e "packed-fp32-ops" -e "^$" "${OBJS[$i]}.log" >&2 || true
done
[[ $fail == 0 ]] || { echo "build failed" >&2; exit 1; }
"$HIPCC" -shared -fPIC --offload-arch=gfx950 -o libqsv.so.tmp "${OBJS[@]}"
mv libqsv.so.tmp libqsv.so
echo "built $HERE/libqsv.so"
# QRunes front end + SDK facade over the QuantumGates adapter (plain C++ over the C ABI)
g++ -O2 -std=c++14 -Wall -fPIC -shared -I"$HERE/cpp" -I"$HERE/../include" -o libqpanda_sdk.so.tmp \
    cpp/QRunes.cpp cpp/QPandaSDK.cpp cpp/HipQuantumGates.cpp cpp/QuantumGates.cpp \
    -L"$HERE" -lqsv -Wl,-rpath,'$ORIGIN'
mv libqpanda_sdk.so.tmp libqpanda_sdk.so
echo "built $HERE/libqpanda_sdk.so"
# QuantumGates adapter (plain C++ over the C ABI) and its known-answer test driver
g++ -O2 -std=c++14 -Wall -I"$HERE/cpp" -I"$HERE/../include" -o _build/test_hip_quantum_gates \
    "$HERE/../tests/cpp/test_hip_quantum_gates.cpp" cpp/HipQuantumGates.cpp cpp/QuantumGates.cpp \
    -L"$HERE" -lqsv -Wl,-rpath,'$ORIGIN/..'
echo "built $HERE/_build/test_hip_quantum_gates"
