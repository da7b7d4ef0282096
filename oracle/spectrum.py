"""Kinetic-energy spectrum of the reference's statistc, restated with numpy
(test infrastructure only: the checker of the device diagnostics).

BOX/saveload.f90:1111-1300 with the transforms of BOX/mpifft.f90:52-108:
each velocity component is transformed with the forward sign and divided by
n on every axis (so u_hat = FFT(u) / (nx ny nz)); E(ik) sums |u_hat|^2 over
the modes 0 <= kx < nx/2 (the r2c half space without its Nyquist plane,
`tmph(1:lx,:,:)`), all ky, kz, with weight 1/2 on the kx = 0 plane (:1275-1278),
in shells ik = int(sqrt(kx^2 + ky^2 + kz^2) + 0.5) (BOX/gaussian.f90:73),
ik = 1..nek.  Signed wavenumbers follow BOX/para.f90:266-278.
"""
from __future__ import annotations

import numpy as np


def signed_k(n: int) -> np.ndarray:
    i = np.arange(n)
    return np.where(i < n // 2 + 1, i, i - n)


def spectrum(ux, uy, uz, nek: int) -> np.ndarray:
    """ux, uy, uz: arrays of shape (nz, ny, nx) (x fastest)."""
    nz, ny, nx = ux.shape
    kz = signed_k(nz)[:, None, None]
    ky = signed_k(ny)[None, :, None]
    kx = np.arange(nx // 2)[None, None, :]
    ik = np.floor(np.sqrt((kx * kx + ky * ky + kz * kz).astype(np.float64)) + 0.5).astype(np.int64)
    w = np.where(kx == 0, 0.5, 1.0)
    ek = np.zeros(nek)
    for u in (ux, uy, uz):
        uh = np.fft.fftn(u) / u.size
        e = (np.abs(uh[:, :, : nx // 2]) ** 2) * w
        for s in range(1, nek + 1):
            ek[s - 1] += float(e[ik == s].sum())
    return ek
