#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/liboracle.so.
# IEEE binary64 with no FMA contraction so every expression rounds as written.
set -e
here=$(cd "$(dirname "$0")" && pwd)
gcc -O3 -march=x86-64-v2 -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared -std=c11 -Wall \
    -o "$here/liboracle.so.tmp" "$here/alens_oracle.c" -lm
mv "$here/liboracle.so.tmp" "$here/liboracle.so"
