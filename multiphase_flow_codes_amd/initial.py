"""Initial-condition generators for the benchmark and test workloads.

``solenoidal_field`` draws the configs' random solenoidal velocity field:
Fourier modes 0 < |k| <= kmax with energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2),
random phases, divergence-free by projection, scaled to a given u_rms
(sqrt of the mean of u.u/3).  The reference itself starts its runs from rest
(BOX/main.f90:73-75); this generator is the workload definition, not a
reference routine.
"""
from __future__ import annotations

import numpy as np


def solenoidal_field(shape, k0=4.0, kmax=16.0, u_rms=0.03, seed=54321):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    kz = np.fft.fftfreq(nz, 1.0 / nz)[:, None, None]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None]
    kx = np.fft.rfftfreq(nx, 1.0 / nx)[None, None, :]
    k2 = kz * kz + ky * ky + kx * kx
    k = np.sqrt(k2)
    mask = (k > 0) & (k <= kmax)
    # |u_hat(k)|^2 * 4 pi k^2 ~ E(k)  ->  |u_hat| ~ k exp(-(k/k0)^2)
    amp = np.where(mask, k * np.exp(-(k / k0) ** 2), 0.0)
    comps = []
    for _ in range(3):
        re = rng.standard_normal(k.shape)
        im = rng.standard_normal(k.shape)
        comps.append((re + 1j * im) * amp)
    ax, ay, az = comps
    kdot = (kx * ax + ky * ay + kz * az) / np.where(k2 > 0, k2, 1.0)
    ax = ax - kx * kdot
    ay = ay - ky * kdot
    az = az - kz * kdot
    u = [np.fft.irfftn(a, s=shape) for a in (ax, ay, az)]
    rms = np.sqrt(sum(np.mean(c * c) for c in u) / 3.0)
    scale = u_rms / rms if rms > 0 else 0.0
    return tuple(np.ascontiguousarray(c * scale) for c in u)
