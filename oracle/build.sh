#!/bin/sh
# Builds the CPU oracle (test infrastructure) as oracle/liboracle.so.
# FP contraction is off and no fast-math flags are used, so FP32 results follow IEEE
# single-precision arithmetic operation by operation. OpenMP threads only the CPU timing of a full
# iteration (oracle_set_threads); the default of one thread keeps every loop sequential.
set -e
cd "$(dirname "$0")"
gcc -std=c11 -O2 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math -Wall -Wextra -o liboracle.so.tmp oracle.c -lm
mv liboracle.so.tmp liboracle.so
