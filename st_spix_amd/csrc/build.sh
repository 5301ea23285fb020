#!/bin/bash
# Builds libstspix.so for gfx950 (MI355X) in-tree.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../_lib"
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wno-unused-command-line-argument \
  -o "$OUT/libstspix.so.tmp" \
  "$HERE/k_bass.hip" "$HERE/k_batch.hip" "$HERE/k_prop.hip" "$HERE/k_color.hip" "$HERE/k_steps.hip" "$HERE/host.hip"
mv "$OUT/libstspix.so.tmp" "$OUT/libstspix.so"
