"""The C-ABI library loads and exports every entry point include/mfc_lbm.h
declares (no device calls: runs without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfc_lbm.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(mfc_[a-z_0-9]+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mfc_create" in names and "mfc_step" in names and "mfc_beads_collision" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol():
    from multiphase_flow_codes_amd import lbm
    if not os.path.exists(lbm.LIB_PATH):
        pytest.fail("libmfc_lbm.so is not built (python -m multiphase_flow_codes_amd.build)")
    L = lbm.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared surface
    assert set(lbm.SIGNATURES) == set(declared_functions())


def test_create_without_device_reports_nodevice():
    """On a host without a GPU, mfc_create fails cleanly (no CPU fallback)."""
    import ctypes
    from multiphase_flow_codes_amd import lbm
    from multiphase_flow_codes_amd.params import LBMParams
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    p = lbm.make_params(LBMParams(nx=8, ny=8, nz=8, visc=0.01))
    h = ctypes.c_void_p()
    rc = lbm.lib().mfc_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == 7   # MFC_ERR_NODEVICE
