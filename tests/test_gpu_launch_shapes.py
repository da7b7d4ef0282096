"""Launch shapes the small parity cases do not reach, against the FP64 oracle.

Rows longer than 256 nodes are split into several x-chunks, one workgroup
each (the 1024^3 channel has four per row). The collide's tile order over
the XCDs (MFC_XCD_REMAP 0, 1 or 2; chosen by shape when unset), the split
of a bounce-back segment over several workgroups (MFC_BB_SPLIT) and the z
planes per box segment (MFC_SEG_PLANES 8; 4 unless the spheres are many)
only change which workgroup computes which nodes or links. Every setting must match the
oracle to 1e-12 in FP64: the split changes the order of the force sums, so
1e-12 rather than bitwise. The tile orders must also agree bitwise with each
other, in FP32 (both arithmetic forms) and in FP64 in-place storage.

Box 544 x 18 x 19: three x-chunks per row (256, 256 and a ragged 32), and
342 rows, so both remaps also have tiles left over after their last whole
group. The spheres sit across the chunk edge at x = 256, across the
periodic x seam, and inside the short last chunk.
"""
import numpy as np
import pytest

from helpers import small_params, start_both

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
N = (544, 18, 19)
POS = np.array([[255.6, 7.3, 8.1], [0.8, 3.2, 12.5], [530.2, 11.4, 4.4]])
POS_WALLS = np.array([[255.6, 7.3, 8.1], [0.8, 7.9, 12.5], [530.2, 7.0, 4.4]])


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _oracle_case(monkeypatch, env, nsteps=4):
    from multiphase_flow_codes_amd.lbm import Lattice
    from oracle.pyoracle import Oracle
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    prm = small_params(n=N, npart=3, rad=3.5, rhopart=1.5)
    lat = Lattice(prm, "fp64")
    lat.prt_init(prm.npart)
    orc = Oracle(prm, bb_mode=1)
    start_both(prm, lat, orc, np.random.default_rng(11), amp=0.02)
    orc.macrovar()
    lat.macrovar()
    orc.ypglb[:] = POS
    orc.yp[:] = POS
    orc.S.istep00 = 1
    orc.call("initpartvel")
    orc.call("initpartomg")
    orc.beads_links()
    lat.prt_set("ypglb", POS)
    lat.initpartvel()
    lat.initpartomg()
    lat.beads_links()
    for it in range(1, nsteps + 1):
        orc.S.istep = it
        orc.call("timestep")
    lat.step(1, nsteps)
    lat.sync()
    return lat, orc


@pytest.mark.parametrize("env", [
    {"MFC_XCD_REMAP": "0"},
    {"MFC_XCD_REMAP": "1"},
    {"MFC_XCD_REMAP": "2"},
    {"MFC_BB_SPLIT": "2"},
    {"MFC_BB_SPLIT": "4"},
    {"MFC_SEG_PLANES": "8"},
    {"MFC_SEG_PLANES": "8", "MFC_BB_SPLIT": "2"},
], ids=["remap0", "remap1", "remap2", "bbsplit2", "bbsplit4", "seg8", "seg8bbsplit2"])
def test_wide_rows_match_oracle(oracle_lib, monkeypatch, env):
    lat, orc = _oracle_case(monkeypatch, env)
    assert orc.S.nconflict == 0
    assert lat.status()["nlinks"] == int((orc.ilink > 0).sum())
    assert rel(lat.download_f(), orc.f) < TOL64
    for name in ("ypglb", "wp", "omgp", "fHIp", "torqp"):
        assert rel(lat.prt_get(name), getattr(orc, name)) < TOL64, name
    ib, isn = lat.get_maps(0)
    assert np.array_equal(ib, orc.ibnodes)
    assert np.array_equal(isn, orc.isnodes)
    rho, ux, uy, uz = lat.download_macro()
    for a, b in ((rho, orc.rho), (ux, orc.ux), (uy, orc.uy), (uz, orc.uz)):
        assert rel(a, b) < TOL64


def _run_remap(monkeypatch, remap, precision, storage, walls):
    from multiphase_flow_codes_amd.lbm import Lattice
    from helpers import smooth_velocity
    monkeypatch.setenv("MFC_XCD_REMAP", remap)
    if walls:
        prm = small_params(n=N, npart=3, rad=3.5, rhopart=2.0, walls=True, forcing=False,
                           strong_forcing=False, body_force=(2e-5, 0.0, 0.0), mingap_w=3.0)
        pos = POS_WALLS
    else:
        prm = small_params(n=N, npart=3, rad=3.5, rhopart=1.5)
        pos = POS
    lat = Lattice(prm, precision, storage=storage)
    ux, uy, uz = smooth_velocity(prm, np.random.default_rng(5), amp=0.02)
    lat.upload_macro(None, ux, uy, uz)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    lat.prt_init(prm.npart)
    lat.prt_set("ypglb", pos)
    lat.initpartvel()
    lat.initpartomg()
    lat.beads_links()
    lat.step(1, 5)   # odd count: in place, the field ends in the neighbours' slots
    lat.sync()
    return lat.download_f(), lat.prt_get("fHIp"), lat.prt_get("ypglb")


@pytest.mark.parametrize("precision,storage,walls", [
    ("fp32", "ab", False),
    ("fp32", "ab", True),
    ("fp32x", "ab", False),
    ("fp64", "aa", True),
], ids=["fp32-periodic", "fp32-channel", "fp32x-periodic", "fp64-aa-channel"])
def test_tile_orders_bitwise_equal(monkeypatch, precision, storage, walls):
    runs = [_run_remap(monkeypatch, r, precision, storage, walls) for r in ("0", "1", "2")]
    for k, other in enumerate(runs[1:], start=1):
        for a, b, what in zip(runs[0], other, ("populations", "fHIp", "ypglb")):
            assert np.array_equal(a, b), f"remap {k} vs 0: {what}"
