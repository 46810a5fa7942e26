#!/usr/bin/env bash
# Build the reference's own CPU demosaic / black-level path from its sources
# where they lie, into oracle/_ref/libimop_ref.so.  Test infrastructure only.
#
# Flags follow the reference build (CMakeLists.txt:30-46, 222-240): C++17,
# -O3, -mavx2 -mfma, OpenMP.  GCC's default for C++ is -ffp-contract=fast, so
# multiply-adds are fused into FMAs exactly as in the reference build.
# -march=native and -flto are left out so the library also runs on other x86
# hosts; neither changes results without -ffast-math.
#
# A second library, libimop_ref_isp.so, holds the CPU white balance, colour
# matrix, gamma, polynomial lens-shading and adaptive defect-correction
# stages, and the raw-data analyzer.  Their module dispatch files (needed for the input
# validation they call) also reference the CUDA entry points, which a build
# without CUDA does not define; compiling with hidden visibility and one
# section per function and linking with --gc-sections drops the unused
# dispatchers and with them those references, so only the stages the shim
# calls remain.
#
# Nothing is copied out of the reference tree: the compiler reads the sources
# in place and only the shared object is written, under oracle/_ref/.
set -euo pipefail
REF=${IMOP_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src/algorithms" ]; then
    echo "build_ref: reference sources not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
SRCS=(
    "$REF/src/algorithms/dispatch.cpp"
    "$REF/src/algorithms/avx2.cpp"
    "$REF/src/algorithms/hqli.cpp"
    "$REF/src/algorithms/mg.cpp"
    "$REF/src/algorithms/prism.cpp"
    "$REF/src/algorithms/super_fast.cpp"
    "$REF/src/algorithms/l7.cpp"
    "$REF/src/algorithms/dfpd.cpp"
    "$REF/src/algorithms/ahd.cpp"
    "$REF/src/algorithms/amaze.cpp"
    "$REF/src/algorithms/rcd.cpp"
    "$REF/src/algorithms/cuda_dispatch.cpp"
    "$REF/src/black_level/global.cpp"
    "$REF/src/black_level/per_channel.cpp"
    "$HERE/ref_shim.cpp"
)
g++ -std=c++17 -O3 -mavx2 -mfma -fopenmp -fPIC -shared -w \
    -I"$REF/include" -I"$REF/src/algorithms" \
    "${SRCS[@]}" -o "$OUT/libimop_ref.so.tmp"
mv "$OUT/libimop_ref.so.tmp" "$OUT/libimop_ref.so"
echo "build_ref: wrote $OUT/libimop_ref.so"

ISP_SRCS=(
    "$REF/src/white_balance/manual.cpp"
    "$REF/src/white_balance/dispatch.cpp"
    "$REF/src/ccm/linear_3x3.cpp"
    "$REF/src/ccm/dispatch.cpp"
    "$REF/src/tone/gamma.cpp"
    "$REF/src/tone/dispatch.cpp"
    "$REF/src/lens_shading/polynomial.cpp"
    "$REF/src/lens_shading/dispatch.cpp"
    "$REF/src/defect_correct/adaptive.cpp"
    "$REF/src/defect_correct/dispatch.cpp"
    "$REF/src/analyzer.cpp"
    "$HERE/ref_isp_shim.cpp"
)
g++ -std=c++17 -O3 -mavx2 -mfma -fPIC -shared -w -fvisibility=hidden -ffunction-sections -fdata-sections \
    -Wl,--gc-sections -I"$REF/include" -I"$REF/src" \
    "${ISP_SRCS[@]}" -o "$OUT/libimop_ref_isp.so.tmp"
mv "$OUT/libimop_ref_isp.so.tmp" "$OUT/libimop_ref_isp.so"
echo "build_ref: wrote $OUT/libimop_ref_isp.so"
