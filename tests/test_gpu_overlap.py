"""The overlapped step (particle phase of step n on a second stream beside
the collide of step n + 1 on the tiles far from every sphere, from the
phase's start with MFC_OVERLAP=1 or after its bounce-back with 2) gives
bitwise the results of the plain step sequence (MFC_OVERLAP=0): populations,
particle state and node maps, for both storages, both precisions, periodic
boxes and walls, across several mfc_step calls."""
import os

import numpy as np
import pytest

from helpers import place_spheres, small_params, smooth_velocity

pytestmark = pytest.mark.gpu


def _make(prm, prec, storage, overlap, pos, u):
    from multiphase_flow_codes_amd.lbm import Lattice
    old = os.environ.get("MFC_OVERLAP")
    os.environ["MFC_OVERLAP"] = str(int(overlap))
    try:
        lat = Lattice(prm, prec, storage=storage)
    finally:
        if old is None:
            del os.environ["MFC_OVERLAP"]
        else:
            os.environ["MFC_OVERLAP"] = old
    lat.upload_macro(None, *u)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    lat.prt_init(prm.npart)
    lat.prt_set("ypglb", pos)
    lat.initpartvel()
    lat.initpartomg()
    lat.beads_links()
    return lat


CASES = [
    ((96, 80, 72), False, "fp64", "ab", 3, 4.0),
    ((100, 64, 48), True, "fp32", "ab", 3, 3.5),
    ((96, 80, 72), False, "fp64", "aa", 3, 4.0),
    ((130, 48, 40), True, "fp32x", "aa", 2, 3.0),
    ((64, 64, 64), False, "fp32", "ab", 8, 4.0),
]


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("n,walls,prec,storage,npart,rad", CASES)
def test_overlapped_step_is_bitwise_the_plain_step(oracle_lib, n, walls, prec, storage, npart, rad, mode):
    rng = np.random.default_rng(sum(n) + npart)
    prm = small_params(n=n, npart=npart, rad=rad, walls=walls, rhopart=1.5)
    u = smooth_velocity(prm, rng, amp=0.03)
    pos = place_spheres(prm, rng, npart, gap=2.5, margin=rad + 3.0 if walls else 0.0)
    plain = _make(prm, prec, storage, 0, pos, u)
    over = _make(prm, prec, storage, mode, pos, u)
    try:
        it = 1
        for chunk in (1, 5, 7):
            plain.step(it, chunk)
            over.step(it, chunk)
            it += chunk
            plain.sync()
            over.sync()
            assert np.array_equal(plain.download_f(), over.download_f()), it
            for name in ("ypglb", "wp", "omgp", "fHIp", "torqp", "flubp"):
                assert np.array_equal(plain.prt_get(name), over.prt_get(name)), (it, name)
            for a, b in zip(plain.get_maps(0), over.get_maps(0)):
                assert np.array_equal(a, b), it
        sp, so = plain.status(), over.status()
        assert sp["nlinks"] == so["nlinks"] > 0 and sp["refill"] == so["refill"]
    finally:
        plain.close()
        over.close()


def test_overlapped_step_timing_counts_steps(oracle_lib):
    """Timing of an overlapped step brackets its two collide launches and
    counts them as one collide."""
    rng = np.random.default_rng(3)
    prm = small_params(n=(96, 64, 64), npart=2, rad=4.0, rhopart=1.5)
    u = smooth_velocity(prm, rng, amp=0.02)
    pos = place_spheres(prm, rng, 2, gap=2.5)
    lat = _make(prm, "fp32", "ab", 1, pos, u)
    try:
        lat.set_timing(1)
        lat.step(1, 6)
        ms, n = lat.get_timing()
        assert n == 6 and ms > 0.0
    finally:
        lat.close()
