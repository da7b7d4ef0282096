"""Many spheres (more than the 512 above which the particle pair searches use
a cell grid instead of all pairs, and more than 4096 box segments, above
which the link offsets are scanned over several workgroups): lubrication
forces, the fused step, node maps and links, and the overlap check against the
FP64 oracle, which always searches all pairs."""
import numpy as np
import pytest

from helpers import small_params, start_both
from multiphase_flow_codes_amd.initial import random_spheres

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _case(walls):
    kw = dict(walls=True, forcing=False, strong_forcing=False, body_force=(1e-5, 0.0, 0.0)) if walls else {}
    prm = small_params(n=(96, 88, 104), npart=1100, rad=2.5, rhopart=1.5, **kw)
    rng = np.random.default_rng(17)
    # surface gaps from 0.4 up: many pairs inside the lubrication range 2r + mingap
    pos = random_spheres(prm.npart, prm.rad, (prm.nz, prm.ny, prm.nx), gap=0.4,
                         wall_clear=1.5 if walls else 0.0, walls_y=walls, seed=17)
    return prm, pos, rng


def _released(prm, pos, rng):
    from multiphase_flow_codes_amd.lbm import Lattice
    from oracle.pyoracle import Oracle
    lat = Lattice(prm, "fp64")
    lat.prt_init(prm.npart)
    orc = Oracle(prm, bb_mode=1)
    start_both(prm, lat, orc, rng, amp=0.01)
    orc.macrovar()
    lat.macrovar()
    orc.ypglb[:] = pos
    orc.yp[:] = pos
    orc.S.istep00 = 1
    orc.call("initpartvel")
    orc.call("initpartomg")
    orc.beads_links()
    lat.prt_set("ypglb", pos)
    lat.initpartvel()
    lat.initpartomg()
    lat.beads_links()
    return lat, orc


@pytest.mark.parametrize("walls", [False, True])
def test_lubrication_from_cells_matches_oracle(oracle_lib, walls):
    prm, pos, rng = _case(walls)
    lat, orc = _released(prm, pos, rng)
    orc.call("beads_lubforce")
    lat.beads_lubforce()
    fo, fg = orc.flubp, lat.prt_get("flubp")
    assert np.count_nonzero(np.abs(fo).sum(axis=1)) > 50      # many spheres feel a partner
    assert rel(fg, fo) < TOL64


@pytest.mark.parametrize("walls", [False, True])
def test_fused_steps_with_cells_match_oracle(oracle_lib, walls):
    prm, pos, rng = _case(walls)
    lat, orc = _released(prm, pos, rng)
    for it in range(1, 4):
        orc.S.istep = it
        orc.call("timestep")
    lat.step(1, 3)
    lat.sync()
    ib, isn = lat.get_maps(0)
    assert np.array_equal(ib, orc.ibnodes) and np.array_equal(isn, orc.isnodes)
    il, ml, al = lat.get_links()
    assert np.array_equal(il, orc.ilink) and np.array_equal(ml, orc.mlink)
    # gaps under a node put boundary links on nodes solid for a neighbour
    # (contact links): both sides take them from the post-collision snapshot
    if not walls:
        assert orc.S.nconflict > 0 and lat.status()["contact_links"] > 0
    assert rel(lat.download_f(), orc.f) < TOL64
    for name in ("ypglb", "wp", "omgp", "fHIp", "torqp", "flubp"):
        assert rel(lat.prt_get(name), getattr(orc, name)) < TOL64, name


def test_overlap_found_through_cells():
    from multiphase_flow_codes_amd.lbm import MfcError
    prm, pos, rng = _case(False)
    pos = pos.copy()
    pos[123] = pos[456] + np.array([2.0 * prm.rad - 0.5, 0.0, 0.0])   # overlapping pair
    pos[123, 0] %= prm.nx
    from multiphase_flow_codes_amd.lbm import Lattice
    lat = Lattice(prm, "fp64")
    lat.prt_init(prm.npart)
    lat.prt_set("ypglb", pos)
    with pytest.raises(MfcError, match="overlap at istep = 5"):
        lat.beads_redistribute(5)     # synchronises, as the reference's fatal check


@pytest.mark.parametrize("bad", ["beyond_wall", "nan"])
def test_sphere_outside_the_box_is_a_fault(bad):
    """A centre beyond a channel wall or not finite: the links scan skips
    the sphere (its box check) and the next synchronisation reports the
    fault, as the reference stops on a sphere it cannot place."""
    from multiphase_flow_codes_amd.lbm import Lattice, MfcError
    from helpers import small_params
    prm = small_params(n=(32, 24, 28), npart=2, rad=3.0, walls=True, forcing=False, strong_forcing=False)
    lat = Lattice(prm, "fp64")
    z = np.zeros((prm.nz, prm.ny, prm.nx))
    lat.upload_macro(None, z, z, z)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    lat.prt_init(prm.npart)
    pos = np.array([[10.3, 12.1, 9.7], [20.5, prm.ny + 5.0, 17.2]])
    if bad == "nan":
        pos[1] = [20.5, 12.0, np.nan]
    lat.prt_set("ypglb", pos)
    try:
        with pytest.raises(MfcError, match="outside the box"):
            lat.beads_links()
            lat.sync()
    finally:
        lat.close()
