#!/bin/bash
# Builds libstspix.so for gfx950 (MI355X) in-tree: one hipcc per source file,
# all in parallel, then one link.  With arguments (source names without
# .hip) only those are recompiled and the other objects are reused.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../_lib"
OBJ="$OUT/obj"
mkdir -p "$OBJ"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC
  -ffp-contract=off -fno-fast-math
  -Xclang -target-feature -Xclang -packed-fp32-ops
  -Wno-unused-command-line-argument)
SRCS=(k_sweep k_split k_batch k_bass k_prop k_color k_steps host)
BUILD=("${SRCS[@]}")
[ $# -gt 0 ] && BUILD=("$@")
pids=()
for s in "${BUILD[@]}"; do
  "$HIPCC" "${FLAGS[@]}" -c "$HERE/$s.hip" -o "$OBJ/$s.o" 2> "$OBJ/$s.err" &
  pids+=($!)
done
rc=0
for i in "${!pids[@]}"; do
  if ! wait "${pids[$i]}"; then
    echo "build: ${BUILD[$i]}.hip failed" >&2
    cat "$OBJ/${BUILD[$i]}.err" >&2
    rc=1
  fi
done
[ $rc -eq 0 ] || exit 1
objs=()
for s in "${SRCS[@]}"; do objs+=("$OBJ/$s.o"); done
"$HIPCC" --offload-arch=gfx950 -fPIC -shared -o "$OUT/libstspix.so.tmp" "${objs[@]}"
mv "$OUT/libstspix.so.tmp" "$OUT/libstspix.so"
