#!/bin/bash
# The host-side planners under AddressSanitizer + UndefinedBehaviorSanitizer:
# builds the library with its host code instrumented (device code as usual)
# into tests/cpp/_san/, builds plan_fuzz against it and runs it on the CPU.
# No device is used.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
SAN="$HERE/_san"
mkdir -p "$SAN"
cd "$ROOT"
python3 -c "
import sys; sys.path.insert(0, '$ROOT')
from painty_amd import build as B
B.build(out='$SAN/libpainty_hip.so', objdir='$SAN/obj',
        extra=['-O1', '-g', '-Xarch_host', '-fsanitize=address,undefined',
               '-Xarch_host', '-fno-sanitize-recover=undefined',
               '-Xarch_host', '-fno-omit-frame-pointer'])
"
HIPCC=/opt/rocm/bin/hipcc
$HIPCC --offload-arch=gfx950 -std=c++17 -O1 -g -Xarch_host -fsanitize=address,undefined \
  -Xarch_host -fno-sanitize-recover=undefined -Xarch_host -fno-omit-frame-pointer \
  -I"$ROOT/include" -I"$ROOT/painty_amd/csrc" -c "$HERE/plan_fuzz.cpp" -o "$SAN/plan_fuzz.o"
/opt/rocm/llvm/bin/clang++ -fsanitize=address,undefined "$SAN/plan_fuzz.o" -o "$SAN/plan_fuzz" \
  -L"$SAN" -lpainty_hip -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,"$SAN" -Wl,-rpath,/opt/rocm/lib
ASAN_OPTIONS="detect_leaks=0:abort_on_error=1" UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1" \
  "$SAN/plan_fuzz" "$@"
