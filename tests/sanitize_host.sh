#!/bin/bash
# Host-side sanitizer run (CPU only, no GPU needed): builds the oracle and the
# C-ABI library with AddressSanitizer + UndefinedBehaviorSanitizer on host code
# (device code is compiled as usual) into a temporary directory, then runs the
# CPU test files that exercise them -- the oracle's pins, the C-ABI argument
# checks and the CNN oracle -- against those builds.
#
#     bash tests/sanitize_host.sh
set -e
R="$(cd "$(dirname "$0")/.." && pwd)"
T="$(mktemp -d)"
trap 'rm -rf "$T"' EXIT
CL=/opt/rocm/llvm/bin/clang
RT="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
$CL -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -march=x86-64 -ffp-contract=off \
    -fPIC -std=c11 -shared -o "$T/liboracle.so" "$R/oracle/oracle.c" -lm -lpthread
F="--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -fvisibility=hidden -ffp-contract=off"
F="$F -fhip-fp32-correctly-rounded-divide-sqrt -Xclang -target-feature -Xclang -packed-fp32-ops"
F="$F -Wno-unused-command-line-argument -I $R/include"
F="$F -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer"
for s in "$R"/imageproc_amd/csrc/*.hip; do
    /opt/rocm/bin/hipcc $F -c -o "$T/$(basename "$s").o" "$s" 2>/dev/null &
done
wait
/opt/rocm/bin/hipcc --offload-arch=gfx950 -shared -fPIC -Xarch_host -fsanitize=address \
    -Xarch_host -fsanitize=undefined -o "$T/libimageproc_hip.so" "$T"/*.hip.o
cd "$R"
LD_PRELOAD="$RT" ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 \
UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 SAN_DIR="$T" python - <<'PY'
import os, sys
sys.path.insert(0, os.getcwd())
import oracle, imageproc_amd._lib as L
oracle._LIB_PATH = os.path.join(os.environ["SAN_DIR"], "liboracle.so")
L.LIB_PATH = os.path.join(os.environ["SAN_DIR"], "libimageproc_hip.so")
import pytest
rc = pytest.main(["-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", "tests/test_oracle.py",
                  "tests/test_capi.py", "tests/test_oracle_cnn.py"])
maps = open("/proc/self/maps").read()
assert oracle._LIB_PATH in maps and L.LIB_PATH in maps, "sanitizer builds were not the ones loaded"
sys.exit(rc)
PY
