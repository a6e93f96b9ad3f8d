#!/bin/bash
# Builds the drop-in C++ demo against libofeli_hip.so (host compiler only).
set -e
ROOT="$(cd "$(dirname "$0")/../.." && pwd)"
g++ -std=c++17 -O2 -Wall -I"$ROOT/ofeli_amd/cxx" -I"$ROOT/include" \
  "$ROOT/tests/cpp/dropin_demo.cpp" -L"$ROOT/ofeli_amd" -lofeli_hip \
  -Wl,-rpath,'$ORIGIN/../../ofeli_amd' -o "$ROOT/tests/cpp/dropin_demo"
