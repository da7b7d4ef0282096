"""Device FP64 diagnostics against numpy restatements: field sums (<= 1e-13
relative), the statistc kinetic-energy spectrum (<= 1e-12 relative to its
largest shell) on power-of-two and other lengths, and after particle steps
(sums of the virtual macrovar fields equal sums of the downloaded arrays)."""
import numpy as np
import pytest

from helpers import place_spheres, small_params, smooth_velocity

pytestmark = pytest.mark.gpu


def _sums_np(rho, ux, uy, uz):
    q = ux * ux + uy * uy + uz * uz
    return np.array([rho.sum(), ux.sum(), uy.sum(), uz.sum(), q.sum(), q.max()])


@pytest.mark.parametrize("n", [(24, 20, 16), (32, 32, 32), (64, 16, 48)])
def test_sums_and_spectrum_match_numpy(n):
    from multiphase_flow_codes_amd.lbm import Lattice
    from oracle.spectrum import spectrum
    prm = small_params(n=n)
    rng = np.random.default_rng(sum(n))
    ux, uy, uz = smooth_velocity(prm, rng, amp=0.03, kmax=4)
    ux = ux + rng.normal(0, 1e-3, ux.shape)   # content at every wavenumber
    rho = rng.normal(0, 1e-3, ux.shape)
    lat = Lattice(prm)
    lat.upload_macro(rho, ux, uy, uz)
    s = lat.diag_sums()
    want = _sums_np(rho, ux, uy, uz)
    got = np.array([s["rho"], s["ux"], s["uy"], s["uz"], s["u2"], s["u2max"]])
    scale = np.array([np.abs(rho).sum(), np.abs(ux).sum(), np.abs(uy).sum(), np.abs(uz).sum(),
                      want[4], want[5]])
    assert np.all(np.abs(got - want) <= 1e-13 * scale)
    nek = min(n) // 2 + 2
    ek = lat.diag_spectrum(nek)
    ek_np = spectrum(ux, uy, uz, nek)
    assert np.max(np.abs(ek - ek_np)) <= 1e-12 * np.max(ek_np)
    s2 = lat.diag_sums()
    assert s2 == s


def test_sums_after_particle_steps_match_downloaded_fields():
    from multiphase_flow_codes_amd.lbm import Lattice
    rng = np.random.default_rng(3)
    prm = small_params(n=(32, 32, 32), npart=3, rad=4.0, rhopart=1.5)
    lat = Lattice(prm)
    lat.prt_init(prm.npart)
    ux, uy, uz = smooth_velocity(prm, rng, amp=0.02)
    lat.upload_macro(None, ux, uy, uz)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    lat.prt_set("ypglb", place_spheres(prm, rng, prm.npart, gap=2.5))
    lat.initpartvel()
    lat.initpartomg()
    lat.beads_links()
    lat.step(1, 5)
    s = lat.diag_sums()
    rho, ux, uy, uz = lat.download_macro()
    want = _sums_np(rho, ux, uy, uz)
    got = np.array([s["rho"], s["ux"], s["uy"], s["uz"], s["u2"], s["u2max"]])
    scale = np.array([np.abs(rho).sum(), np.abs(ux).sum(), np.abs(uy).sum(), np.abs(uz).sum(),
                      want[4], want[5]])
    assert np.all(np.abs(got - want) <= 1e-13 * scale)
