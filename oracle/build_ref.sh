#!/usr/bin/env bash
# Builds oracle/_ref/ref_driver: the reference's own var_inc.f90, para.f90
# and collision.f90, compiled unchanged from where they lie under
# /root/reference with amdflang (default REAL promoted to 8 bytes, as the
# reference was built), linked with oracle/ref_driver.f90.
#
# The reference's other step routines (streaming.f90, partlib.f90,
# gaussian.f90, initial.f90) need an MPI Fortran module for this compiler
# and/or FFTW 2.x, neither of which this image provides, so they are not
# built; the oracle's restatements of them are pinned by the tests instead.
#
# Outputs go only to oracle/_ref/ (git-ignored).  Exits 0 without building
# when the reference tree or amdflang is absent.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref="/root/reference/3DBox_Forced_HI_Turbulence_Moving_Particles_LBM"
out="$here/_ref"
fc="${FC:-amdflang}"
if [ ! -d "$ref" ] || ! command -v "$fc" >/dev/null 2>&1; then
  echo "build_ref: reference tree or $fc not available; skipping" >&2
  exit 0
fi
mkdir -p "$out/mod"
flags="-fdefault-real-8 -O1 -ffp-contract=off -J $out/mod"
"$fc" $flags -c "$ref/var_inc.f90" -o "$out/var_inc.o"
"$fc" $flags -c "$ref/para.f90" -o "$out/para.o"
"$fc" $flags -c "$ref/collision.f90" -o "$out/collision.o"
"$fc" $flags -c "$here/ref_driver.f90" -o "$out/ref_driver.o"
"$fc" -o "$out/ref_driver" "$out/ref_driver.o" "$out/var_inc.o" "$out/para.o" "$out/collision.o"
echo "build_ref: built $out/ref_driver"
