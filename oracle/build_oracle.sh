#!/bin/sh
# Builds the FP64 oracle (test infrastructure) as oracle/liboracle.so.
# -ffp-contract=off keeps every f32/f64 operation individually rounded, as in
# the reference's Rust code (no implicit fused multiply-add).
set -e
here="$(cd "$(dirname "$0")" && pwd)"
g++ -std=c++17 -O2 -ffp-contract=off -fno-fast-math -fPIC -shared -pthread \
    -Wall -Wextra -Wno-unused-parameter \
    "$here/mpm_oracle.cpp" -o "$here/liboracle.so.tmp"
mv "$here/liboracle.so.tmp" "$here/liboracle.so"
