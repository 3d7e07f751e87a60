#!/bin/bash
# Builds the FP64 oracle with AddressSanitizer and UndefinedBehaviorSanitizer
# and runs the CPU test suite (pytest -m "not gpu") against that build.
# Python itself is not instrumented, so the sanitizer runtimes are preloaded;
# leak checking is off (the interpreter keeps its allocations at exit).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(dirname "$HERE")"
make -s -C "$HERE" sanitize
ASAN="$(gcc -print-file-name=libasan.so)"
UBSAN="$(gcc -print-file-name=libubsan.so)"
export PAINTY_ORACLE_LIB="$HERE/_san/liboracle.so"
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:allocator_may_return_null=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
cd "$ROOT"
LD_PRELOAD="$ASAN $UBSAN" python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
