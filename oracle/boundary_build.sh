#!/usr/bin/env bash
# TEST INFRASTRUCTURE ONLY.  Builds oracle/_ref/libhip_boundary.so: the C++
# HIPAccelerator (libraw_enhanced_amd/csrc/hip_accelerator.cpp) compiled
# against the reference's own core/accelerator.h, next to the reference's
# CPUAccelerator (its object from ref_build.sh), behind the facade selection of
# tests/boundary_shim.cpp.  Reads the reference headers where they lie; needs
# ref_build.sh to have run and liblre_hip.so to be built.
set -euo pipefail
REF="${LRE_REFERENCE_ROOT:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
OUT="$HERE/_ref"
LIB="$ROOT/libraw_enhanced_amd/_lib"
if [ ! -f "$REF/core/accelerator.h" ] || [ ! -f "$OUT/cpu_accelerator.o" ] || [ ! -f "$LIB/liblre_hip.so" ]; then
    echo "boundary_build: reference tree, oracle/_ref or liblre_hip.so missing; skipping" >&2
    exit 0
fi
CXX="${CXX_REF:-/opt/rocm/lib/llvm/bin/clang++}"
FLAGS=(-std=c++17 -O2 -fno-fast-math -ffp-contract=off -fopenmp -fPIC -w -I"$REF/core" -I"$ROOT/include"
       -I"$ROOT/libraw_enhanced_amd/csrc")
"$CXX" "${FLAGS[@]}" -c "$ROOT/libraw_enhanced_amd/csrc/hip_accelerator.cpp" -o "$OUT/hip_accelerator.o"
"$CXX" "${FLAGS[@]}" -c "$ROOT/tests/boundary_shim.cpp" -o "$OUT/boundary_shim.o"
"$CXX" -shared -fopenmp -o "$OUT/libhip_boundary.so" "$OUT/boundary_shim.o" "$OUT/hip_accelerator.o" \
    "$OUT/cpu_accelerator.o" -L"$LIB" -llre_hip -Wl,-rpath,"\$ORIGIN/../../libraw_enhanced_amd/_lib" \
    -Wl,-rpath,/opt/rocm/lib/llvm/lib -L/opt/rocm/lib/llvm/lib
echo "boundary_build: $OUT/libhip_boundary.so"
