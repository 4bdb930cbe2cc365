#!/bin/sh
# Builds the CPU oracle (test infrastructure) next to its source: oracle/libmpm_oracle.so
set -e
here="$(cd "$(dirname "$0")" && pwd)"
g++ -O3 -march=x86-64-v3 -fopenmp -shared -fPIC -std=c++17 -Wall -Wno-unknown-pragmas \
    -o "$here/libmpm_oracle.so.tmp" "$here/cpu_ref.cpp"
mv "$here/libmpm_oracle.so.tmp" "$here/libmpm_oracle.so"
