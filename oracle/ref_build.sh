#!/usr/bin/env bash
# TEST INFRASTRUCTURE ONLY.  Builds the reference's own CPU implementation of
# the develop path into oracle/_ref/libref_cpu.so, straight from the sources
# where they lie in the read-only reference tree.  Nothing is copied into this
# repository; the objects and the library go to oracle/_ref/ (git-ignored).
#
#  * core/cpu_accelerator.cpp is compiled whole except lines 3070-3319
#    (enhance_micro_contrast and its Gaussian helper), the one block that
#    needs Apple's <simd/simd.h>; no stage of the develop path uses it.  The
#    file also relies on <cstring> arriving transitively on Apple, so it is
#    force-included here.
#  * core/libraw_wrapper.cpp cannot be compiled as a unit (it needs LibRaw and
#    pybind11).  Its detail tone map is compiled from its own lines:
#    box_filter_normalized (lines 45-116, a free function with its pragma) and
#    the two Impl member functions aces_tone_map_scalar and
#    apply_detail_preserving_tonemap (lines 1389-1471), placed inside a struct
#    so the member functions have a class to live in.
#
# The build is IEEE-strict (-O2 -fno-fast-math -ffp-contract=off, no -march);
# the reference's own Apple build uses -Ofast, so Apple binaries may differ in
# the last bits.
set -euo pipefail
REF="${LRE_REFERENCE_ROOT:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
if [ ! -f "$REF/core/cpu_accelerator.cpp" ]; then
    echo "ref_build: reference tree not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
CXX="${CXX_REF:-/opt/rocm/lib/llvm/bin/clang++}"
FLAGS=(-std=c++17 -O2 -fno-fast-math -ffp-contract=off -fopenmp -fPIC -w -I"$REF/core")

sed -e '3070,3319d' "$REF/core/cpu_accelerator.cpp" |
    "$CXX" -x c++ "${FLAGS[@]}" -include cstring -c - -o "$OUT/cpu_accelerator.o"

{
    echo '#include "accelerator.h"'
    echo '#include <algorithm>'
    echo '#include <iostream>'
    echo '#include <vector>'
    echo 'namespace libraw_enhanced {'
    sed -n '45,116p' "$REF/core/libraw_wrapper.cpp"
    echo 'struct DetailToneMapRef {'
    sed -n '1389,1471p' "$REF/core/libraw_wrapper.cpp"
    echo '};'
    echo '}'
    echo 'extern "C" void ref_detail_tm_impl(float *rgb, int w, int h) {'
    echo '  libraw_enhanced::ImageBufferFloat b;'
    echo '  b.image = reinterpret_cast<float (*)[3]>(rgb); b.width = w; b.height = h; b.channels = 3;'
    echo '  libraw_enhanced::DetailToneMapRef().apply_detail_preserving_tonemap(b);'
    echo '}'
} | "$CXX" -x c++ "${FLAGS[@]}" -c - -o "$OUT/detail_tonemap.o"

"$CXX" "${FLAGS[@]}" -c "$HERE/ref_harness.cpp" -o "$OUT/ref_harness.o"
"$CXX" -shared -fopenmp -o "$OUT/libref_cpu.so" "$OUT/cpu_accelerator.o" \
    "$OUT/detail_tonemap.o" "$OUT/ref_harness.o" \
    -Wl,-rpath,/opt/rocm/lib/llvm/lib -L/opt/rocm/lib/llvm/lib
echo "ref_build: $OUT/libref_cpu.so"
