"""Pins of the numpy spectrum restatement (oracle/spectrum.py) by closed
forms: a single Fourier mode of amplitude A carries A^2/4 in its shell
(half-space mode with kx > 0, or the conjugate pair on the kx = 0 plane with
weight 1/2 each), and a field without Nyquist content satisfies Parseval:
sum_ik E(ik) = <|u|^2>/2 when all its modes lie in shells 1..nek."""
import numpy as np

from oracle.spectrum import spectrum


def _grid(nz, ny, nx):
    return np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")


def test_single_mode_kx_positive():
    nz, ny, nx = 12, 10, 16
    z, y, x = _grid(nz, ny, nx)
    A = 0.37
    u = A * np.cos(2 * np.pi * (3 * x / nx + 2 * y / ny + 0.3))
    ek = spectrum(u, 0 * u, 0 * u, 6)
    shell = int(np.sqrt(9 + 4) + 0.5)
    want = np.zeros(6)
    want[shell - 1] = A * A / 4
    assert np.allclose(ek, want, rtol=0, atol=1e-15)


def test_single_mode_on_kx0_plane():
    nz, ny, nx = 12, 10, 16
    z, y, x = _grid(nz, ny, nx)
    A = 0.21
    u = A * np.sin(2 * np.pi * (1 * y / ny - 2 * z / nz))
    ek = spectrum(0 * u, u, 0 * u, 5)
    want = np.zeros(5)
    want[int(np.sqrt(5) + 0.5) - 1] = A * A / 4
    assert np.allclose(ek, want, rtol=0, atol=1e-15)


def test_parseval_band_limited():
    rng = np.random.default_rng(0)
    nz = ny = nx = 16
    z, y, x = _grid(nz, ny, nx)
    comps = []
    for _ in range(3):
        u = np.zeros((nz, ny, nx))
        for _m in range(6):
            k = rng.integers(-3, 4, size=3)
            u += rng.normal() * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) / nx + rng.uniform(0, 6))
        comps.append(u - u.mean())
    ek = spectrum(*comps, 8)
    assert abs(ek.sum() - 0.5 * sum((c * c).mean() for c in comps)) < 1e-13
