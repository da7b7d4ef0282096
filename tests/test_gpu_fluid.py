"""GPU parity of the fluid step (no particles) against the FP64 oracle.

Bar: the FP64 library (the drop-in) matches the oracle to <= 1e-12 normwise
relative error per step; the FP32-shifted path is compared with the FP64 path
by the relative L2 norm of the velocity field.
"""
import numpy as np
import pytest

from helpers import fluid_step_both, small_params, start_both, smooth_velocity

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _pair(prm, precision="fp64", bb_mode=1):
    from multiphase_flow_codes_amd.lbm import Lattice
    from oracle.pyoracle import Oracle
    return Lattice(prm, precision), Oracle(prm, bb_mode=bb_mode)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("walls", [False, True])
def test_fp64_fluid_steps_match_oracle(oracle_lib, walls):
    prm = small_params(walls=walls)
    lat, orc = _pair(prm)
    rng = np.random.default_rng(3)
    start_both(prm, lat, orc, rng)
    assert rel(lat.download_f(), orc.f) < TOL64
    for step in range(6):
        fluid_step_both(lat, orc)
        f = lat.download_f()
        assert rel(f, orc.f) < TOL64, f"step {step}: f"
        rho, ux, uy, uz = lat.download_macro()
        for a, b in ((rho, orc.rho), (ux, orc.ux), (uy, orc.uy), (uz, orc.uz)):
            assert rel(a, b) < TOL64, f"step {step}: macro"
        fx, fy, fz = lat.download_force()
        for a, b in ((fx, orc.fx), (fy, orc.fy), (fz, orc.fz)):
            assert rel(a, b) < 1e-13, f"step {step}: force field"
    assert np.max(np.abs(orc.fx)) > 0.0  # forcing was active


def test_prerelaxation_array_mode_matches_oracle(oracle_lib):
    """Pre-relaxation loop (BOX/main.f90:86-108): collision with the frozen
    velocity arrays and the density from rhoupdat, before any FORCING."""
    prm = small_params(forcing=False)
    lat, orc = _pair(prm)
    rng = np.random.default_rng(5)
    start_both(prm, lat, orc, rng, amp=0.03)
    for it in range(5):
        rhop = orc.rho.copy()
        orc.collision_mrt()
        orc.streaming()
        orc.rhoupdat()
        err_o = np.max(np.abs(orc.rho - rhop))
        lat.collision_mrt()
        lat.streaming()
        err_g = lat.rhoupdat()
        assert abs(err_g - err_o) <= 1e-12 * max(err_o, 1e-300) + 1e-18, f"iteration {it}"
        assert rel(lat.download_f(), orc.f) < TOL64


def test_prerelax_helper_converges_like_the_loop(oracle_lib):
    """Lattice.prerelax runs the same loop to the residual threshold: the
    iteration count and field equal the oracle's loop with the same exit."""
    prm = small_params(forcing=False)
    lat, orc = _pair(prm)
    rng = np.random.default_rng(6)
    start_both(prm, lat, orc, rng, amp=0.03)
    eps = 1e-7
    it_o = 0
    while True:
        rhop = orc.rho.copy()
        orc.collision_mrt()
        orc.streaming()
        orc.rhoupdat()
        err_o = np.max(np.abs(orc.rho - rhop))
        if err_o <= eps or it_o > 15000:
            break
        it_o += 1
    it_g, err_g = lat.prerelax(eps)
    assert it_g == it_o and it_o > 3
    assert abs(err_g - err_o) <= 1e-6 * err_o
    assert rel(lat.download_f(), orc.f) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32x"])
def test_fused_step_equals_fine_grained(oracle_lib, precision):
    """mfc_step, split into two calls, equals FORCING, collision, streaming,
    macrovar called one by one: populations, forcing field and forcing
    process state, bitwise."""
    from multiphase_flow_codes_amd.lbm import Lattice
    prm = small_params()
    rng = np.random.default_rng(9)
    a, b = Lattice(prm, precision), Lattice(prm, precision)
    ux, uy, uz = smooth_velocity(prm, rng)
    for lat in (a, b):
        lat.upload_macro(None, ux, uy, uz)
        lat.initpop()
        lat.rhoupdat()
        lat.macrovar()
    for it in range(1, 6):
        a.forcing(); a.collision_mrt(); a.streaming(); a.macrovar()
    b.step(1, 3)
    b.step(4, 2)
    b.sync()
    assert np.array_equal(a.download_f(), b.download_f())
    for x, y in zip(a.download_force(), b.download_force()):
        assert np.array_equal(x, y)
    fa, fb = a.get_forcing_state(), b.get_forcing_state()
    assert (fa["iseedf"], fa["iyf"]) == (fb["iseedf"], fb["iyf"])
    assert np.array_equal(fa["ivf"], fb["ivf"])
    for key in ("a", "b"):
        for x, y in zip(fa[key], fb[key]):
            assert np.array_equal(x, y), key


def test_upload_download_roundtrip_both_phases(oracle_lib):
    from multiphase_flow_codes_amd.lbm import Lattice
    prm = small_params(walls=True)
    lat = Lattice(prm)
    rng = np.random.default_rng(1)
    f = rng.normal(0, 1e-3, (prm.nz, prm.ny, prm.nx, 19))
    lat.upload_f(f)                      # post-streaming phase
    assert np.array_equal(lat.download_f(), f)
    lat.upload_macro(np.zeros_like(f[..., 0]), *smooth_velocity(prm, rng))
    lat.collision_mrt()                  # now pre-streaming
    g = lat.download_f()
    lat.upload_f(g)
    assert np.array_equal(lat.download_f(), g)


@pytest.mark.parametrize("precision", ["fp32", "fp32x"])
def test_fp32_tracks_fp64(oracle_lib, precision):
    """FP32 storage (FP32 or FP64 collision arithmetic) stays within 1e-5
    velocity relative L2 of the FP64 path over 200 forced steps."""
    from multiphase_flow_codes_amd.lbm import Lattice
    prm = small_params(n=(32, 32, 32), strong_forcing=False)
    rng = np.random.default_rng(2)
    ux, uy, uz = smooth_velocity(prm, rng, amp=0.03)
    lats = {p: Lattice(prm, p) for p in ("fp64", precision)}
    for lat in lats.values():
        lat.upload_macro(None, ux, uy, uz)
        lat.initpop()
        lat.rhoupdat()
        lat.macrovar()
        lat.step(1, 200)
    u64 = np.stack(lats["fp64"].download_macro()[1:])
    u32 = np.stack(lats[precision].download_macro()[1:])
    assert np.all(np.isfinite(u64))
    err = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    assert err < 1e-5, err


def test_sampled_collide_timing(oracle_lib):
    """mfc_set_timing(every=N) brackets launches 0, N, 2N, ... with events."""
    from multiphase_flow_codes_amd.lbm import Lattice
    prm = small_params(n=(32, 32, 32))
    lat = Lattice(prm)
    lat.upload_macro(None, *smooth_velocity(prm, np.random.default_rng(4)))
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    lat.set_timing(3)
    lat.step(1, 10)
    ms, n = lat.get_timing()
    assert n == 4 and ms > 0.0
    lat.set_timing(0)
    lat.step(11, 2)
    assert lat.get_timing() == (0.0, 0)
