#!/bin/sh
# Builds the reference decoder from its own sources, where they lie in the reference checkout, into oracle/_ref/:
#   libref_scalar.so  gcc -DOB_SCALAR: the scalar form (the one research/SPEC.md states and wasm runs)
#   libref_x86.so     ROCm clang -march=x86-64-v2: the native x86-64 SIMD form (liblizard/core/CMakeLists.txt:160-167)
#   libref_xfer.so    the transfer's BLAKE3 tree (xfer.c) over the vendored BLAKE3 with its SSE4.1 / AVX2 kernels, and the
#                     test stream's SHAKE256 (shake.c)
# Both with -ffp-contract=off as every reference target (liblizard/build.sh:15-17). Test infrastructure only.
# Usage: oracle/ref/build_ref.sh [reference root, default /root/reference]
set -e
REF=${1:-/root/reference}
SRC="$REF/liblizard/src"
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/../_ref"
if [ ! -f "$SRC/lizard_phy.c" ]; then echo "build_ref: no reference sources at $SRC" >&2; exit 2; fi
mkdir -p "$OUT"
UNITS="$SRC/ldpc.c $SRC/layout.c $HERE/ref_acquire.c $SRC/fft.c $SRC/fmt.c $SRC/rs.c"  # acquire.c through ref_acquire.c
gcc -O2 -ffp-contract=off -DOB_SCALAR -fPIC -shared -I"$SRC" -o "$OUT/libref_scalar.so.tmp" "$HERE/ref_driver.c" $UNITS -lm
mv "$OUT/libref_scalar.so.tmp" "$OUT/libref_scalar.so"
CLANG=/opt/rocm/llvm/bin/clang
if [ -x "$CLANG" ]; then
  "$CLANG" -O2 -ffp-contract=off -march=x86-64-v2 -fPIC -shared -I"$SRC" -o "$OUT/libref_x86.so.tmp" "$HERE/ref_driver.c" $UNITS -lm
  mv "$OUT/libref_x86.so.tmp" "$OUT/libref_x86.so"
fi
# the transfer's BLAKE3 tree (liblizard/src/xfer.c) over the vendored official BLAKE3 C, its SSE4.1 and AVX2 kernels
# (the unix assembly; the SSE2 and AVX-512 ones are not vendored), as the reference's native build links it
B3="$REF/liblizard/vendor/blake3"
gcc -O2 -fPIC -shared -Wl,-z,defs -DBLAKE3_NO_SSE2 -DBLAKE3_NO_AVX512 -I"$SRC" -I"$B3" -o "$OUT/libref_xfer.so.tmp" "$SRC/xfer.c" "$SRC/shake.c" "$B3/blake3.c" \
  "$B3/blake3_dispatch.c" "$B3/blake3_portable.c" "$B3/blake3_sse41_x86-64_unix.S" "$B3/blake3_avx2_x86-64_unix.S"
mv "$OUT/libref_xfer.so.tmp" "$OUT/libref_xfer.so"
echo "build_ref: $(ls "$OUT")"
