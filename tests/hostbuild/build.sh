#!/bin/bash
# Host build of the library for checks on the CPU (test infrastructure, not a product path).
#
# The HIP sources of power-grid-model_amd/csrc are compiled as plain C++ with g++ against hip/hip_runtime.h of
# this folder, which runs every kernel launch on host threads (one thread per GPU thread of a block, the blocks
# one after another), keeps streams synchronous and replays graphs in capture order (host_runtime.cpp). The build
# uses AddressSanitizer and UndefinedBehaviorSanitizer, and the arena takes one allocation per request so every
# device array has its own bounds. Run the GPU tests against it with ./run.sh (EMU_MMAP=1: device memory in an
# address range apart from host memory, with pointer and copy-direction checks; EMU_FILL=<byte>: fill of fresh
# device memory; EMU_DEFER=1: streams run only when waited on, so missing waits between streams show; EMU_DEFER=2:
# a destroyed stream's work runs only at the next device-wide synchronisation). Output: tests/hostbuild/_build/libpower_grid_model_c.so.
set -e
cd "$(dirname "$0")"
rm -rf _build && mkdir -p _build/src _build/obj
cp ../../power-grid-model_amd/csrc/* _build/src/
# POOL=1 keeps the library's own arena (256 MiB chunks) and chunk pool instead
[ "${POOL:-0}" = 1 ] || python3 - <<'PY'
p='_build/src/engine.cpp'
s=open(p).read()
a=s.index('void* Arena::take(size_t bytes) {')
b=s.index('\n}\n', a)+3
s=s[:a]+'''void* Arena::take(size_t bytes) {
    if (bytes == 0) return nullptr;
    void* p = nullptr;
    PGM_HIP_CHECK(hipMalloc(&p, bytes));
    chunks_.emplace_back(p, bytes);
    reserved_ += bytes;
    return p;
}
'''+s[b:]
old='if (pool_)\n            pool_->put(c.first, c.second);\n        else\n            (void)hipFree(c.first);'
assert old in s
s=s.replace(old,'(void)hipFree(c.first);')
open(p,'w').write(s)
PY
SAN="-fsanitize=address,undefined -fno-sanitize=vptr"
FLAGS="-std=c++20 -O1 -g -fPIC -fno-omit-frame-pointer $SAN -I. -Wno-unused-result -Wno-attributes -Wno-unknown-pragmas"
pids=()
for f in _build/src/*.cpp _build/src/*.hip; do
  g++ -x c++ $FLAGS -c "$f" -o "_build/obj/$(basename "$f").o" &
  pids+=($!)
done
g++ $FLAGS -c host_runtime.cpp -o _build/obj/host_runtime.o &
pids+=($!)
for p in "${pids[@]}"; do wait "$p"; done
g++ -shared $SAN -o _build/libpower_grid_model_c.so _build/obj/*.o -lpthread
echo built
