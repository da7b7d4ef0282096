"""The unchanged driver's call sequence (BOX/main.f90:73-188) through the
Fortran drop-in wrappers.

* Against the FP64 oracle: populations, macroscopic fields and particle state
  (with the integrator history that savecntdpart writes) to <= 1e-12
  normwise relative error after the pre-relaxation, the release at step 2 and
  six steps with moving spheres; the node maps handed back to var_inc
  (ibnodes, isnodes and the previous pair) bit-exact; the fluid rms velocity
  of rmsstat (BOX/saveload.f90:1553-1591, masked by ibnodes) to 1e-12; the
  forcing random-number state equal.
* A run stopped after step 5, restarted from only what the reference's
  restart files hold (a fresh device context) and continued to step 9 ends
  bitwise where the uninterrupted run ends.  The forcing is off in this case:
  with it on, the reference's own restart differs from the uninterrupted run
  (force_real* are not in the files, so the first macrovar after the restart
  has no forcing term).
* MFC_PRECISION / MFC_STORAGE select the precision and the storage without a
  source change: in-place FP64 storage is bitwise the two-buffer run, and the
  FP32 run restarts bitwise too.
"""
import dataclasses
import os

import numpy as np
import pytest

from fortran_dropin import DRIVER, PART_OUT, inputs, oracle_run, read_outputs, run_driver, write_inputs

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built: run tests/fortran/build.sh")


def _run(tmp, prm, u, nsteps, nrelax, nsplit=0, env=None):
    os.makedirs(tmp, exist_ok=True)
    write_inputs(tmp, prm, u, nsteps=nsteps, nrelax=nrelax, nsplit=nsplit)
    r = run_driver(tmp, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    return read_outputs(tmp, prm)


def rms_fluid(ux, uy, uz, ibnodes):
    """velfrms of rmsstat: rms of the fluctuation about the fluid mean over
    nodes with ibnodes < 0."""
    m = ibnodes < 0
    nf = int(m.sum())
    tot = 0.0
    for c in (ux, uy, uz):
        mn = c[m].sum() / nf
        tot += ((c[m] - mn) ** 2).sum() / nf
    return np.sqrt(tot)


def test_fortran_dropin_matches_oracle(oracle_lib, tmp_path):
    _need_driver()
    prm, u = inputs()
    nsteps, nrelax = 8, 3
    got = _run(str(tmp_path), prm, u, nsteps, nrelax)
    orc, it = oracle_run(prm, u, nsteps, nrelax)
    assert got["nrelax"] == it
    assert orc.S.nconflict == 0
    assert rel(got["f"], orc.f) < TOL64
    for name in ("rho", "ux", "uy", "uz"):
        assert rel(got[name], getattr(orc, name)) < TOL64, name
    for name in PART_OUT:
        b = getattr(orc, name)
        if np.max(np.abs(b)) == 0.0:
            assert np.max(np.abs(got[name])) == 0.0, name
        else:
            assert rel(got[name], b) < TOL64, name
    for name in ("ibnodes", "isnodes", "ibnodes0", "isnodes0"):
        assert np.array_equal(got[name], getattr(orc, name)), name
    # the spheres moved: the two maps differ, and both hold all three spheres
    assert not np.array_equal(got["ibnodes"], got["ibnodes0"]) or not np.array_equal(got["isnodes"], got["isnodes0"])
    assert set(np.unique(got["isnodes"])) == {-1, 1, 2, 3}
    a = rms_fluid(got["ux"], got["uy"], got["uz"], got["ibnodes"])
    b = rms_fluid(orc.ux, orc.uy, orc.uz, orc.ibnodes)
    assert abs(a - b) <= 1e-12 * b
    assert got["iseedf"] == orc.S.iseedf
    assert got["iyf"] == orc.S.iyf


def _unforced():
    prm, u = inputs(seed=11)
    return dataclasses.replace(prm, vscale=0.0), u


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fortran_dropin_restart_is_bitwise(tmp_path, prec):
    _need_driver()
    prm, u = _unforced()
    env = {"MFC_PRECISION": prec}
    whole = _run(str(tmp_path / "whole"), prm, u, 9, 2, env=env)
    split = _run(str(tmp_path / "split"), prm, u, 9, 2, nsplit=5, env=env)
    for name in ("f", "rho", "ux", "uy", "uz") + PART_OUT:
        assert np.array_equal(whole[name], split[name]), name
    for name in ("ibnodes", "isnodes", "ibnodes0", "isnodes0"):
        assert np.array_equal(whole[name], split[name]), name
    assert (whole["iseedf"], whole["iyf"]) == (split["iseedf"], split["iyf"])
    # the particle history is live: the spheres moved and accelerate
    assert np.max(np.abs(whole["dwdt"])) > 0.0 and np.max(np.abs(whole["wp"])) > 0.0


def test_fortran_dropin_storage_from_environment(tmp_path):
    _need_driver()
    prm, u = inputs()
    ab = _run(str(tmp_path / "ab"), prm, u, 6, 2, env={"MFC_STORAGE": "ab"})
    aa = _run(str(tmp_path / "aa"), prm, u, 6, 2, env={"MFC_STORAGE": "aa"})
    for name in ("f", "rho", "ux", "uy", "uz") + PART_OUT:
        assert np.array_equal(ab[name], aa[name]), name
    bad = run_driver(str(tmp_path / "ab"), env={"MFC_PRECISION": "fp16"})
    assert bad.returncode != 0 and "MFC_PRECISION" in bad.stdout + bad.stderr
