"""The ensemble engine (mfc_ens_*): boxes of different sizes, sphere counts,
radii, densities and body forces advanced together with one grid per step
stage give bitwise the same populations, particle state and node maps as the
same boxes advanced one at a time with mfc_step, including a box that joins
late and one that leaves early, and for channel boxes with walls; the collide
block and tile-order autotuner keeps results unchanged, and so does replaying
the steps as recorded HIP graphs."""
import numpy as np
import pytest

from helpers import place_spheres, small_params, smooth_velocity

pytestmark = pytest.mark.gpu

CASES = [
    # (n, npart, rad, rhopart, body force x, seed)
    ((24, 20, 28), 2, 3.0, 2.5, 2e-5, 1),
    ((32, 32, 32), 3, 4.0, 5.0, 5e-6, 2),
    ((40, 36, 24), 1, 3.5, 1.5, 1e-4, 3),
    ((28, 28, 28), 0, 3.0, 5.0, 3e-5, 4),
]


def _box(case, precision):
    from multiphase_flow_codes_amd.lbm import Lattice
    n, npart, rad, rhop, bf, seed = case
    prm = small_params(n=n, npart=npart, rad=rad, rhopart=rhop, forcing=False, visc=0.01,
                       strong_forcing=False, body_force=(bf, 0.0, 0.0))
    rng = np.random.default_rng(seed)
    lat = Lattice(prm, precision)
    ux, uy, uz = smooth_velocity(prm, rng, amp=0.02)
    lat.upload_macro(None, ux, uy, uz)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    if npart:
        lat.prt_init(npart)
        lat.prt_set("ypglb", place_spheres(prm, rng, npart, gap=2.5))
        lat.initpartvel()
        lat.initpartomg()
        lat.beads_links()
    return lat


def _same(a, b):
    assert np.array_equal(a.download_f(), b.download_f())
    if a.prm.npart:
        for name in ("ypglb", "wp", "omgp", "fHIp", "torqp", "flubp"):
            assert np.array_equal(a.prt_get(name), b.prt_get(name)), name
        for which in (0, 1):
            ia, sa = a.get_maps(which)
            ib, sb = b.get_maps(which)
            assert np.array_equal(ia, ib) and np.array_equal(sa, sb)


@pytest.mark.parametrize("precision,graph", [("fp64", False), ("fp32", False), ("fp32", True),
                                             ("fp32x", True)])
def test_ensemble_equals_boxes_alone(precision, graph):
    from multiphase_flow_codes_amd.lbm import Ensemble
    alone = [_box(c, precision) for c in CASES]
    members = [_box(c, precision) for c in CASES]
    ens = Ensemble()
    ens.set_graph(graph)
    for lat in members[:3]:
        ens.add(lat, 1)
    ens.step(4)
    ens.add(members[3], 1)            # joins late, at its own step 1
    ens.step(3)
    ens.remove(members[0])            # leaves after 7 steps
    ens.step(2)
    ens.sync()
    steps = [7, 9, 9, 5]
    for lat, n in zip(alone, steps):
        lat.step(1, n)
        lat.sync()
    for a, b in zip(alone, members):
        _same(a, b)
    ens.close()


def test_autotune_keeps_results():
    from multiphase_flow_codes_amd.lbm import Ensemble
    alone = [_box(c, "fp32") for c in CASES]
    members = [_box(c, "fp32") for c in CASES]
    ens = Ensemble()
    for lat in members:
        ens.add(lat, 1)
    (bs, order), times = ens.autotune(2)                # 12 x 2 steps
    assert bs in (64, 128, 256) and order in (0, 1)
    assert len(times) == 6 and all(t > 0 for t in times.values())
    ens.step(3)
    ens.sync()
    for lat in alone:
        lat.step(1, 27)
        lat.sync()
    for a, b in zip(alone, members):
        _same(a, b)
    ens.close()


WALL_CASES = [
    # (n, npart, rad, rhopart, body force x, seed): channel boxes, walls at y = 0 and y = ny
    ((32, 28, 24), 2, 3.0, 2.5, 3e-5, 11),
    ((40, 32, 36), 3, 3.5, 5.0, 1e-5, 12),
    ((24, 24, 20), 0, 3.0, 2.5, 5e-5, 13),
]


def _channel_box(case, precision):
    from multiphase_flow_codes_amd.lbm import Lattice
    n, npart, rad, rhop, bf, seed = case
    prm = small_params(n=n, npart=npart, rad=rad, rhopart=rhop, forcing=False, walls=True, visc=0.01,
                       strong_forcing=False, body_force=(bf, 0.0, 0.0))
    rng = np.random.default_rng(seed)
    lat = Lattice(prm, precision)
    ux, uy, uz = smooth_velocity(prm, rng, amp=0.02)
    lat.upload_macro(None, ux, uy, uz)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    if npart:
        lat.prt_init(npart)
        lat.prt_set("ypglb", place_spheres(prm, rng, npart, gap=2.5, margin=rad + 3.0))
        lat.initpartvel()
        lat.initpartomg()
        lat.beads_links()
    return lat


@pytest.mark.parametrize("precision,graph", [("fp64", False), ("fp32", True)])
def test_channel_ensemble_equals_boxes_alone(precision, graph):
    """Boxes with walls (the channel variant) in one ensemble: the batched
    kernels' wall variants give each box bitwise its single-box result."""
    from multiphase_flow_codes_amd.lbm import Ensemble
    alone = [_channel_box(c, precision) for c in WALL_CASES]
    members = [_channel_box(c, precision) for c in WALL_CASES]
    ens = Ensemble()
    ens.set_graph(graph)
    for lat in members:
        ens.add(lat, 1)
    ens.step(11)
    ens.sync()
    for lat in alone:
        lat.step(1, 11)
        lat.sync()
    for a, b in zip(alone, members):
        _same(a, b)
    ens.close()


def test_membership_errors_are_reported():
    """The engine refuses what it cannot batch, with a message and without
    touching the boxes: a second precision, in-place storage, a box twice, a
    box in two ensembles, a box driven by the forcing process, and the
    autotuner on an empty ensemble; a refused box still steps alone."""
    from multiphase_flow_codes_amd.lbm import Ensemble, Lattice, MfcError
    a32 = _box(CASES[0], "fp32")
    b64 = _box(CASES[1], "fp64")
    z = np.zeros((24, 24, 24))

    def ready(lat):
        lat.upload_macro(None, z, z, z)
        lat.initpop()
        lat.rhoupdat()
        lat.macrovar()
        return lat

    aa = ready(Lattice(small_params(n=(24, 24, 24), forcing=False, strong_forcing=False), "fp32",
                       storage="aa"))
    forced = ready(Lattice(small_params(n=(24, 24, 24)), "fp32"))
    ens, other = Ensemble(), Ensemble()
    try:
        with pytest.raises(MfcError, match="empty ensemble"):
            ens.autotune(1)
        ens.add(a32, 1)
        with pytest.raises(MfcError, match="one precision"):
            ens.add(b64, 1)
        with pytest.raises(MfcError, match="already a member"):
            ens.add(a32, 1)
        with pytest.raises(MfcError, match="another ensemble"):
            other.add(a32, 1)
        with pytest.raises(MfcError, match="two-buffer storage"):
            ens.add(aa, 1)
        with pytest.raises(MfcError, match="body force only"):
            ens.add(forced, 1)
        assert len(ens.members) == 1
        ens.step(2)
        ens.sync()
        b64.step(1, 2)       # the refused box steps alone as before
        b64.sync()
        assert np.isfinite(b64.download_f()).all()
    finally:
        ens.close()
        other.close()
        for lat in (a32, b64, aa, forced):
            lat.close()
