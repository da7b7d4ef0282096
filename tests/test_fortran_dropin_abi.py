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
