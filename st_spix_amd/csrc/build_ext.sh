#!/bin/bash
# Builds the thin torch C++ extension st_spix_amd/bist_cuda*.so (host code
# only, against PyTorch-ROCm's headers), linked to libstspix.so, in-tree.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
PKG="$HERE/.."
PY_INC=$(python3 -c "import sysconfig; print(sysconfig.get_paths()['include'])")
TORCH=$(python3 -c "import os, torch; print(os.path.dirname(torch.__file__))")
EXT=$(python3 -c "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))")
ABI=$(python3 -c "import torch; print(int(torch._C._GLIBCXX_USE_CXX11_ABI))")
ROCM="${ROCM_PATH:-/opt/rocm}"
OUT="$PKG/bist_cuda$EXT"
g++ -O2 -std=c++17 -fPIC -shared -w \
  -D__HIP_PLATFORM_AMD__=1 -DUSE_ROCM=1 -DTORCH_EXTENSION_NAME=bist_cuda \
  -DTORCH_API_INCLUDE_EXTENSION_H -D_GLIBCXX_USE_CXX11_ABI=$ABI \
  -I"$TORCH/include" -I"$TORCH/include/torch/csrc/api/include" -I"$PY_INC" -I"$ROCM/include" \
  "$HERE/bist_ext.cpp" -o "$OUT.tmp" \
  -L"$TORCH/lib" -lc10 -lc10_hip -ltorch -ltorch_cpu -ltorch_hip -ltorch_python \
  -L"$PKG/_lib" -lstspix -Wl,-rpath,'$ORIGIN/_lib'
mv "$OUT.tmp" "$OUT"
