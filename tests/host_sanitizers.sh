#!/bin/bash
# The CPU test suite under AddressSanitizer + UndefinedBehaviorSanitizer, host code only, in two
# passes (a sanitizer runtime has to be the first library of its process, and the two
# compilers bring different ones):
#   1. the native libraries' host code (argument checks, ownership of the arrays handed to the
#      caller, destructors, the host ground state, caches), built by hipcc with the sanitizers
#      on the host compilation alone, with clang's runtime preloaded;
#   2. the C oracle, built by gcc with its own runtime preloaded.
# Nothing here runs on a GPU.  Usage: bash tests/host_sanitizers.sh [OUT_DIR]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=${1:-/tmp/mma_amd_sanitize}
mkdir -p "$OUT/oracle"
cd "$ROOT"

echo "building the host-sanitized native libraries into $OUT/lib"
python -m mma_amd.build --sanitize "$OUT" > "$OUT/build.log" 2>&1
echo "building the sanitized oracle into $OUT/oracle"
gcc -O2 -g -ffp-contract=off -fPIC -std=gnu99 -fsanitize=address,undefined -fno-sanitize-recover=undefined \
    -fno-omit-frame-pointer -fopenmp -shared -o "$OUT/oracle/libctdse_oracle.so" oracle/ctdse_oracle.c -lm

COMMON="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1"
UB="print_stacktrace=1:halt_on_error=1"

echo "pass 1: native libraries (clang runtime)"
env LD_PRELOAD="$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)" \
    LD_LIBRARY_PATH=/opt/rocm/llvm/lib ASAN_OPTIONS=$COMMON UBSAN_OPTIONS=$UB MMA_AMD_LIB_DIR="$OUT/lib" \
    python -m pytest tests -q -m "not gpu" -p no:cacheprovider

echo "pass 2: oracle (gcc runtime)"
env LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)" \
    ASAN_OPTIONS=$COMMON UBSAN_OPTIONS=$UB ORACLE_LIB_DIR="$OUT/oracle" \
    python -m pytest tests -q -m "not gpu" -p no:cacheprovider
