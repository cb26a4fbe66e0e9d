#!/bin/bash
# pytest against the host build: tests/hostbuild/run.sh <pytest arguments>, e.g.
#   tests/hostbuild/run.sh tests/test_gpu_parity.py -m gpu -x -q -k test_validation_case_gpu
HERE="$(cd "$(dirname "$0")" && pwd)"
export PGM_AMD_LIB="$HERE/_build/libpower_grid_model_c.so"
export LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1:print_summary=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
cd "$HERE/../.."
exec python -m pytest "$@"
