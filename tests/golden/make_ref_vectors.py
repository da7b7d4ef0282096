"""Writes tests/golden/ref_slab_vectors.npz: outputs of the compiled reference
(oracle/_ref/ref_driver, built by oracle/build_ref.sh from the reference's
unchanged var_inc.f90, para.f90 and collision.f90) on small slab states, so
the oracle's collision_MRT, rhoupdat, macrovar and the para constants stay
pinned to the reference where that build is absent.

Inputs are not stored: tests/helpers.py rebuilds them bit for bit from the
seeds and shapes kept in the fixture (hash_field).  Run from the repository
root after oracle/build_ref.sh:  python tests/golden/make_ref_vectors.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

from helpers import PARA_NAMES, hashed_state, run_ref_driver, slab_params, sphere_slab  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
OUT = os.path.join(HERE, "ref_slab_vectors.npz")

# (name, mode, ly, lz, seed); mode as in run_ref_driver
CASES = {
    "collision": (1, 2, 3, 0),
    "rhoupdat": (3, 2, 3, 5),
    "macrovar": (2, 4, 4, 11),
    "para": (4, 1, 1, 0),
}
SPHERES = np.array([[20.3, 2.1, 1.9], [64.7, 1.8, 2.2], [101.2, 2.2, 2.05]])


def state(name):
    mode, ly, lz, seed = CASES[name]
    npart = len(SPHERES) if name == "macrovar" else 0
    o = Oracle(slab_params(ly, lz, npart=npart))
    hashed_state(o, seed, solid_frac=0.0 if name == "macrovar" else 0.1)
    if name == "macrovar":
        sphere_slab(o, seed, SPHERES)
    return o


def main():
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} is not built (oracle/build_ref.sh)")
    out = {"spheres": SPHERES}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (mode, ly, lz, seed) in CASES.items():
            o = state(name)
            extra = {}
            if mode == 2:
                extra = dict(isn=o.isnodes, ypglb=o.ypglb, wp=o.wp, omgp=o.omgp)
            f, rho, ux, uy, uz, scal = run_ref_driver(DRIVER, tmp, mode, ly, lz, o.f, o.rho, o.ux,
                                                      o.uy, o.uz, o.fx, o.fy, o.fz, o.ibnodes, **extra)
            out[f"{name}_case"] = np.array([mode, ly, lz, seed], dtype=np.int64)
            if name == "collision":
                out["collision_f"] = f
            elif name == "rhoupdat":
                out["rhoupdat_rho"] = rho
            elif name == "macrovar":
                for k, a in zip(("rho", "ux", "uy", "uz"), (rho, ux, uy, uz)):
                    out[f"macrovar_{k}"] = a
            else:
                out["para"] = scal[:len(PARA_NAMES)]
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
