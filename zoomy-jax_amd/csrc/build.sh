#!/usr/bin/env bash
# Builds libzoomy_hip.so (HIP kernels for gfx950 + host C ABI) in-tree.
# Usage: build.sh [output_dir]   (default: the package directory)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${1:-$HERE/../zoomy_hip}"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
ARCH="${ZJ_OFFLOAD_ARCH:-gfx950}"
mkdir -p "$HERE/_build"
# Packed FP32 instructions are disabled for all device code.
DEVFLAGS=(-O3 -std=c++17 -fPIC --offload-arch="$ARCH"
          -Xclang -target-feature -Xclang -packed-fp32-ops
          -munsafe-fp-atomics -ffp-contract=fast)
OBJS=()
PIDS=()
for src in "$HERE"/*.hip; do
    obj="$HERE/_build/$(basename "${src%.hip}").o"
    "$HIPCC" "${DEVFLAGS[@]}" -c "$src" -o "$obj" &
    PIDS+=($!)
    OBJS+=("$obj")
done
"$HIPCC" -O2 -std=c++17 -fPIC -c "$HERE/zj_rcm.cpp" -o "$HERE/_build/zj_rcm.o"
"$HIPCC" -O2 -std=c++17 -fPIC -c "$HERE/zj_patches.cpp" -o "$HERE/_build/zj_patches.o"
"$HIPCC" -O2 -std=c++17 -fPIC -c "$HERE/zj_io.cpp" -o "$HERE/_build/zj_io.o"
for pid in "${PIDS[@]}"; do wait "$pid"; done
OBJS+=("$HERE/_build/zj_rcm.o" "$HERE/_build/zj_patches.o" "$HERE/_build/zj_io.o")
"$HIPCC" -shared -fPIC --offload-arch="$ARCH" -o "$OUT/libzoomy_hip.so.tmp" "${OBJS[@]}" -ldl -lpthread
mv -f "$OUT/libzoomy_hip.so.tmp" "$OUT/libzoomy_hip.so"
echo "built $OUT/libzoomy_hip.so"
