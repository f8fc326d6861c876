#!/bin/sh
# Builds the oracle (test infrastructure) into oracle/liborc.so, and the same
# source at -O3 for AVX2-class x86-64 hosts (x86-64-v3: Intel Haswell and AMD
# Zen 2 onwards) into oracle/liborc_fast.so, the CPU baseline's build (the
# host of the GPU is not known when this runs, so not -march=native).
# No FMA contraction and no fast-math: the restatement keeps the reference's
# floating-point operation order.
set -e
cd "$(dirname "$0")"
FLAGS="-fopenmp -fPIC -shared -std=c11 -ffp-contract=off -fno-fast-math -Wall -Wextra -Wno-unused-parameter"
gcc -O2 $FLAGS -o liborc.so oracle.c -lm
gcc -O3 -march=x86-64-v3 $FLAGS -o liborc_fast.so oracle.c -lm
