"""The Fortran drop-in (multiphase_flow_codes_amd/fortran/) builds, its
ISO_C_BINDING mirror of struct mfc_params has the C layout, and the driver
reaches the library (CPU: no device, so mfc_create reports MFC_ERR_NODEVICE).
"""
import ctypes
import os
import subprocess

import pytest

from fortran_dropin import DRIVER, LAYOUT, case_params, inputs, write_inputs


def _need(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (tests/fortran/build.sh needs amdflang)")


def test_fortran_params_layout_matches_c():
    _need(LAYOUT)
    from multiphase_flow_codes_amd.lbm import MfcParams
    out = subprocess.run([LAYOUT], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict((ln.split()[0], int(ln.split()[1])) for ln in out if ln.strip())
    assert got.pop("sizeof") == ctypes.sizeof(MfcParams)
    want = {name: getattr(MfcParams, name).offset for name, _ in MfcParams._fields_}
    assert got == want


def test_fortran_driver_reaches_library(tmp_path):
    _need(DRIVER)
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present: the GPU test runs the driver for real")
    except ImportError:
        pass
    prm, u = inputs()
    write_inputs(str(tmp_path), prm, u, nsteps=2, nrelax=2)
    r = subprocess.run([DRIVER], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "fault: mfc_create failed" in r.stdout and "7" in r.stdout


REF_VAR_INC = "/root/reference/3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/var_inc.f90"


def test_dropin_compiles_against_reference_var_inc(tmp_path):
    """The wrappers use only names, kinds and shapes the reference's own
    module var_inc declares: they compile against it unchanged (default REAL
    of 8 bytes, as the reference is built)."""
    import shutil
    fc = os.environ.get("FC", "amdflang")
    if shutil.which(fc) is None:
        pytest.skip(f"{fc} not available")
    if not os.path.exists(REF_VAR_INC):
        pytest.skip("reference tree not present")
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "multiphase_flow_codes_amd", "fortran")
    flags = ["-fdefault-real-8", "-O0", "-J", str(tmp_path)]
    for src in (REF_VAR_INC, os.path.join(pkg, "mfc_lbm_c.f90"), os.path.join(pkg, "lbm_dropin.f90")):
        obj = os.path.join(str(tmp_path), os.path.basename(src) + ".o")
        r = subprocess.run([fc, *flags, "-c", src, "-o", obj], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
