"""The C oracle against committed outputs of the compiled reference
(tests/golden/ref_slab_vectors.npz, written by tests/golden/make_ref_vectors.py
from oracle/_ref/ref_driver: the reference's own var_inc.f90 + para.f90 +
collision.f90 of BOX/ = 3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/).
The inputs are rebuilt from the seeds and shapes in the fixture, so these
checks run without the reference tree or a Fortran compiler:
collision_MRT (BOX/collision.f90:8-210) within one ulp, rhoupdat and
macrovar with solid nodes of three moving spheres within 4 ulp, the para
constants (BOX/para.f90) and their host mirror exactly.  A planted mistake
in the oracle (tests/test_oracle_golden.py::test_fixture_sees_a_planted_mistake)
shows the comparison is not vacuous."""
import os

import numpy as np
import pytest

from helpers import PARA_NAMES, hashed_state, slab_params, sphere_slab

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_slab_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _state(gold, name):
    from oracle.pyoracle import Oracle
    mode, ly, lz, seed = (int(v) for v in gold[f"{name}_case"])
    cen = gold["spheres"] if name == "macrovar" else np.zeros((0, 3))
    o = Oracle(slab_params(ly, lz, npart=len(cen)))
    hashed_state(o, seed, solid_frac=0.0 if name == "macrovar" else 0.1)
    if len(cen):
        sphere_slab(o, seed, cen)
    return o


def _ulps(a, b, n):
    return np.max(np.abs(a - b)) <= 1e-18 + n * 1.1e-16 * np.max(np.abs(b))


def test_collision_mrt_matches_reference_vectors(oracle_lib, gold):
    o = _state(gold, "collision")
    f_in = o.f.copy()
    solid = o.ibnodes > 0
    assert 0 < solid.sum() < solid.size
    o.collision_mrt()
    ref = gold["collision_f"]
    assert ref.shape == o.f.shape
    assert _ulps(o.f, ref, 2)
    assert np.array_equal(ref[solid], f_in[solid])     # the reference skips solid nodes
    assert not np.array_equal(ref[~solid], f_in[~solid])


def test_rhoupdat_matches_reference_vectors(oracle_lib, gold):
    o = _state(gold, "rhoupdat")
    o.rhoupdat()
    assert _ulps(o.rho, gold["rhoupdat_rho"], 4)


def test_macrovar_with_spheres_matches_reference_vectors(oracle_lib, gold):
    o = _state(gold, "macrovar")
    assert (o.ibnodes > 0).sum() >= 30
    o.macrovar()
    for k in ("rho", "ux", "uy", "uz"):
        assert _ulps(getattr(o, k), gold[f"macrovar_{k}"], 4), k


def test_para_constants_match_reference_vectors(gold):
    prm = slab_params(1, 1)
    for name, v in zip(PARA_NAMES, gold["para"]):
        assert getattr(prm, name) == v, name


def test_fixture_sees_a_planted_mistake(oracle_lib, gold):
    """The oracle with two relaxation rates swapped (s1 for s2, energy and
    energy squared: a plausible slip in the collision's diagonal) no longer
    matches the vectors."""
    o = _state(gold, "collision")
    o.P.s1, o.P.s2 = o.P.s2, o.P.s1
    o.collision_mrt()
    assert not _ulps(o.f, gold["collision_f"], 1e6)
