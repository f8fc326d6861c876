#!/bin/sh
# Builds the oracle (test infrastructure) into oracle/liborc.so.
# No FMA contraction and no fast-math: the restatement keeps the reference's
# floating-point operation order.
set -e
cd "$(dirname "$0")"
gcc -O2 -fopenmp -fPIC -shared -std=c11 -ffp-contract=off -fno-fast-math -Wall -Wextra -Wno-unused-parameter \
    -o liborc.so oracle.c -lm
