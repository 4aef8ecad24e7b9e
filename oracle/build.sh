#!/usr/bin/env bash
# Builds the oracle's C restatement (greedy coloring, OpenMP element assembly)
# into oracle/_build/liboracle.so.  Test infrastructure only; the product never
# links it.  No -march: the library runs on whichever x86-64 host the tests or
# the benchmark's CPU baseline run on.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
gcc -O3 -fPIC -shared -std=c99 -fopenmp -o "$here/_build/liboracle.so.tmp" \
    "$here/coloring.c" "$here/fem_omp.c" -lm
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
