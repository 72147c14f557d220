#!/usr/bin/env bash
# Runs the oracle's CPU tests against its sanitizer builds (test
# infrastructure only): AddressSanitizer + UndefinedBehaviorSanitizer with
# GCC, then ThreadSanitizer with LLVM's OpenMP runtime and its Archer tool
# (oracle/tsan.supp lists what is suppressed and why). Each run stops at the
# first report; exit status 0 means both were clean.
set -euo pipefail
cd "$(dirname "$0")/.."
LLVM=${LLVM:-/opt/rocm/llvm}
make -C oracle sanitize
TESTS="tests/test_oracle_pins.py tests/test_oracle_pins_map.py tests/test_oracle_pins_tracking.py tests/test_oracle_track_frame.py"

echo "== ASan + UBSan"
ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 \
UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
LD_PRELOAD="$(g++ -print-file-name=libasan.so):$(g++ -print-file-name=libubsan.so)" \
ORACLE_LIB=oracle/_san/liboracle_asan.so \
    python -m pytest $TESTS -q -x -s -m "not gpu" -p no:cacheprovider

echo "== TSan"
TSAN_RTS=("$LLVM"/lib/clang/*/lib/linux/libclang_rt.tsan-x86_64.so)
TSAN_RT=${TSAN_RTS[0]}
TSAN_OPTIONS="halt_on_error=1 report_signal_unsafe=0 ignore_noninstrumented_modules=1 suppressions=$PWD/oracle/tsan.supp" \
OMP_TOOL_LIBRARIES="$LLVM/lib/libarcher.so" \
LD_PRELOAD="$TSAN_RT" \
ORACLE_LIB=oracle/_san/liboracle_tsan.so \
    python -m pytest $TESTS -q -x -s -m "not gpu" -p no:cacheprovider
