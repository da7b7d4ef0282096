#!/usr/bin/env bash
# Builds tests/fortran/_build/dropin_driver: the Fortran drop-in wrappers
# (multiphase_flow_codes_amd/fortran/) compiled against the test var_inc
# module and linked with libmfc_lbm.so, and layout_check, which prints the
# Fortran layout of struct mfc_params.  Exits 0 without building when
# amdflang is absent.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
pkg="$(cd "$here/../../multiphase_flow_codes_amd" && pwd)"
out="$here/_build"
fc="${FC:-amdflang}"
if ! command -v "$fc" >/dev/null 2>&1; then
  echo "fortran drop-in: $fc not available; skipping" >&2
  exit 0
fi
mkdir -p "$out"
flags="-fdefault-real-8 -O2 -J $out"
"$fc" $flags -c "$here/var_inc.f90" -o "$out/var_inc.o"
"$fc" $flags -c "$pkg/fortran/mfc_lbm_c.f90" -o "$out/mfc_lbm_c.o"
"$fc" $flags -c "$pkg/fortran/lbm_dropin.f90" -o "$out/lbm_dropin.o"
"$fc" $flags -c "$here/dropin_driver.f90" -o "$out/dropin_driver.o"
"$fc" -o "$out/dropin_driver" "$out/dropin_driver.o" "$out/lbm_dropin.o" \
  "$out/mfc_lbm_c.o" "$out/var_inc.o" -L"$pkg" -lmfc_lbm -Wl,-rpath,'$ORIGIN/../../../multiphase_flow_codes_amd'
"$fc" $flags -c "$here/layout_check.f90" -o "$out/layout_check.o"
"$fc" -o "$out/layout_check" "$out/layout_check.o" "$out/mfc_lbm_c.o"
echo "fortran drop-in: built $out/dropin_driver, $out/layout_check"
"$fc" $flags -c "$here/ckpt_io.f90" -o "$out/ckpt_io.o"
"$fc" -o "$out/ckpt_io" "$out/ckpt_io.o"
echo "fortran drop-in: built $out/ckpt_io"
"$fc" $flags -c "$here/big_record.f90" -o "$out/big_record.o"
"$fc" -o "$out/big_record" "$out/big_record.o"
echo "fortran drop-in: built $out/big_record"
