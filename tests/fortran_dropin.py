"""Runs the Fortran drop-in test driver (tests/fortran/dropin_driver.f90) and
the same call sequence on the C oracle.

The driver's constants are fixed by tests/fortran/var_inc.f90 (32^3, three
spheres of radius 4 and density 1.5, release at step 2); the run constants
(MRT rates, forcing scales, seeds) come from LBMParams through a namelist, the
way para would have set them in module var_inc.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

from helpers import small_params, smooth_velocity

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "fortran", "_build")
DRIVER = os.path.join(BUILD, "dropin_driver")
LAYOUT = os.path.join(BUILD, "layout_check")

N = 32
NPART = 3
IRELEASE = 2


def case_params():
    prm = small_params(n=(N, N, N), npart=NPART, rad=4.0, rhopart=1.5, irelease=IRELEASE)
    return prm


def _fmt(v):
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v)) + "d0"


def write_inputs(workdir, prm, u, nsteps, nrelax, rhoepsl=1e-30, nsplit=0):
    vals = {n: getattr(prm, n) for n in ("s1", "s2", "s4", "s9", "s10", "s13", "s16", "omegepsl",
                                         "omegepslj", "omegxx", "ww0", "ww1", "ww2", "volp", "amp",
                                         "aip", "gscale", "rhog", "stf0", "stf1", "vscale",
                                         "tscale", "pi2")}
    vals.update(iseedf=prm.iseedf, iseedp=prm.iseedp, rhoepsl=rhoepsl, nsteps=nsteps,
                nrelax=nrelax, nsplit=nsplit)
    with open(os.path.join(workdir, "dropin_in.nml"), "w") as fh:
        fh.write("&run\n")
        for k, v in vals.items():
            fh.write(f"  {k} = {_fmt(v)},\n")
        fh.write("/\n")
    np.concatenate([np.ascontiguousarray(c).ravel() for c in u]).astype("<f8").tofile(
        os.path.join(workdir, "dropin_u.bin"))


def run_driver(workdir, timeout=120, env=None):
    full = dict(os.environ)
    full.update(env or {})
    return subprocess.run([DRIVER], cwd=workdir, capture_output=True, text=True, timeout=timeout,
                          env=full)


PART_OUT = ("ypglb", "wp", "omgp", "fHIp", "torqp", "flubp",
            "forcepp", "torqpp", "dwdt", "domgdt", "thetap")


def read_outputs(workdir, prm):
    raw = open(os.path.join(workdir, "dropin_out.bin"), "rb").read()
    ncell = prm.ncell
    nd = 19 * ncell + 4 * ncell + len(PART_OUT) * 3 * NPART
    d = np.frombuffer(raw[:8 * nd], dtype="<f8")
    ints = np.frombuffer(raw[8 * nd:], dtype="<i4")
    assert ints.size == 3 + 4 * ncell, ints.size
    shp = (prm.nz, prm.ny, prm.nx)
    out = {"f": d[:19 * ncell].reshape(shp + (19,))}
    o = 19 * ncell
    for name in ("rho", "ux", "uy", "uz"):
        out[name] = d[o:o + ncell].reshape(shp)
        o += ncell
    for name in PART_OUT:
        out[name] = d[o:o + 3 * NPART].reshape(NPART, 3)
        o += 3 * NPART
    out["iseedf"], out["iyf"], out["nrelax"] = (int(x) for x in ints[:3])
    for k, name in enumerate(("ibnodes", "isnodes", "ibnodes0", "isnodes0")):
        out[name] = ints[3 + k * ncell:3 + (k + 1) * ncell].reshape(shp)
    return out


def oracle_run(prm, u, nsteps, nrelax, rhoepsl=1e-30):
    """BOX/main.f90:73-188 on the oracle, in the reference's bounce-back order
    (the drop-in's default)."""
    from oracle.pyoracle import Oracle
    orc = Oracle(prm, bb_mode=0)
    orc.ux[:], orc.uy[:], orc.uz[:] = u
    orc.initpop()
    orc.rhoupdat()
    istep = 0
    while True:
        orc.collision_mrt()
        orc.streaming()
        rhop = orc.rho.copy()
        orc.rhoupdat()
        err = np.max(np.abs(orc.rho - rhop))
        if err <= rhoepsl or istep + 1 >= nrelax:
            break
        istep += 1
    orc.macrovar()
    for it in range(1, nsteps + 1):
        orc.S.istep = it
        if it == prm.irelease:
            orc.release()
        assert orc.call("timestep") == 0
    return orc, istep + 1


def inputs(seed=6):
    prm = case_params()
    rng = np.random.default_rng(seed)
    return prm, smooth_velocity(prm, rng, amp=0.02)
