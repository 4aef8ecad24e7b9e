#!/usr/bin/env bash
# Builds the oracle's C restatement into oracle/_build/liboracle.so.
# Test infrastructure only; the product never links it.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
gcc -O2 -fPIC -shared -std=c99 -o "$here/_build/liboracle.so.tmp" "$here/coloring.c"
mv "$here/_build/liboracle.so.tmp" "$here/_build/liboracle.so"
