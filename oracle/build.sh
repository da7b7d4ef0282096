#!/usr/bin/env bash
# Builds the C oracle (test infrastructure) into oracle/liboracle.so and, when
# the reference tree is present, the reference-collision driver in oracle/_ref.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
${CC:-gcc} -O2 -ffp-contract=off -fopenmp -fPIC -shared -Wall -Wno-unknown-pragmas \
  -o "$here/liboracle.so" "$here/mfc_oracle.c" -lm
bash "$here/build_ref.sh"
