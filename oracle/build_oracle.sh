#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/lib/.
# -ffp-contract=off: no fused multiply-add anywhere in the restatement.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/lib"
gcc -O2 -fPIC -std=c11 -Wall -Wextra -ffp-contract=off -fopenmp -shared \
    -o "$here/lib/libmpcd_oracle.so" "$here/mpcd_oracle.c" "$here/md_oracle.c" -lm
# The CPU baseline (bench.py): the same sources with the reference's flags
# (src/Makefile:2: -O3 -ffast-math -fopenmp); only its OpenMP bucket step
# (orc_omp_*) is used from this one, and only for timing.
gcc -O3 -ffast-math -fPIC -std=c11 -fopenmp -shared \
    -o "$here/lib/libmpcd_baseline.so" "$here/mpcd_oracle.c" "$here/md_oracle.c" -lm
