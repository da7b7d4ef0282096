"""Device synthesis of the initial velocity (mfc_init_velocity) against the
host generator: the same band-limited solenoidal field to rounding, rho = 0,
and initpop from it equal to initpop from the uploaded host field."""
import numpy as np
import pytest

from multiphase_flow_codes_amd.initial import solenoidal_field, solenoidal_table
from multiphase_flow_codes_amd.params import LBMParams

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,k0,walls", [((36, 40, 48), 4.0, False), ((40, 34, 64), 8.0, True)])
def test_device_field_matches_host(shape, k0, walls):
    from multiphase_flow_codes_amd.lbm import Lattice
    nz, ny, nx = shape
    prm = LBMParams(nx=nx, ny=ny, nz=nz, visc=0.01, forcing=False, walls_y=walls)
    host = solenoidal_field(shape, k0=k0, u_rms=0.03, seed=54321)
    table, scale = solenoidal_table(shape, k0=k0, u_rms=0.03, seed=54321)
    a = Lattice(prm, "fp64")
    a.init_velocity(table, scale)
    rho, ux, uy, uz = a.download_macro()
    assert np.all(rho == 0.0)
    top = max(np.abs(c).max() for c in host)
    for d, h in zip((ux, uy, uz), host):
        assert np.max(np.abs(d - h)) <= 1e-13 * top
    # initpop from either field: equal populations to rounding
    b = Lattice(prm, "fp64")
    b.upload_macro(None, *host)
    a.initpop()
    b.initpop()
    fa, fb = a.download_f(), b.download_f()
    assert np.max(np.abs(fa - fb)) <= 1e-13 * np.max(np.abs(fb))
    a.close()
    b.close()
