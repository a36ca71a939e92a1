#!/bin/sh
# Builds the CPU oracle (test infrastructure) as oracle/liboracle.so.
# FP contraction is off and no fast-math flags are used, so FP32 results follow IEEE
# single-precision arithmetic operation by operation.
set -e
cd "$(dirname "$0")"
gcc -std=c11 -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wextra -o liboracle.so.tmp oracle.c -lm
mv liboracle.so.tmp liboracle.so
