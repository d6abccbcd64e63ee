#!/bin/bash
# Builds libanuga_sw.so (the shallow-water step, from sw_capi.hip, and the
# device mesh builder, from mesh_build.hip) and
# libanuga_gnn.so (the learned edge-flux surrogate, from gnn_kernels.hip) for
# gfx950, in-tree.
# -ffp-contract=off: no FMA contraction, so FP64 results follow the C oracle's
#                    operation order bit for bit.
# -packed-fp32-ops disabled: hipcc must not emit packed FP32 instructions.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="${HERE}/../libanuga_sw.so"
ARCH="${SW_OFFLOAD_ARCH:-gfx950}"
/opt/rocm/bin/hipcc --offload-arch="${ARCH}" -O3 -std=c++17 -fPIC -shared \
    -ffp-contract=off -fno-fast-math \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function \
    -o "${OUT}.tmp" "${HERE}/sw_capi.hip" "${HERE}/mesh_build.hip"
mv -f "${OUT}.tmp" "${OUT}"
echo "built ${OUT}"
GOUT="${HERE}/../libanuga_gnn.so"
/opt/rocm/bin/hipcc --offload-arch="${ARCH}" -O3 -std=c++17 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function \
    -o "${GOUT}.tmp" "${HERE}/gnn_kernels.hip"
mv -f "${GOUT}.tmp" "${GOUT}"
echo "built ${GOUT}"
