"""Initial-condition generators for the benchmark and test workloads.

``solenoidal_field`` draws the benchmark workloads' random solenoidal velocity field:
Fourier modes 0 < |k| <= kmax with energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2),
random phases, divergence-free by projection, scaled to a given u_rms
(sqrt of the mean of u.u/3).  The reference itself starts its runs from rest
(BOX/main.f90:73-75); this generator is the workload definition, not a
reference routine.

The field is band-limited: only the modes with |k| <= kmax are drawn (in a
fixed order, from the seed alone), so the same seed gives the same continuous
field on every grid, and it is synthesised separably (x, then y, then z in
slabs of planes), so a 1024^3 field costs seconds and a few GB of host memory.
"""
from __future__ import annotations

import numpy as np


def solenoidal_modes(k0=4.0, kmax=16.0, seed=54321):
    """Integer wavevectors of one half-space (kx > 0, or kx = 0 and ky > 0, or
    kx = ky = 0 and kz > 0) with |k| <= kmax, and their solenoidal complex
    amplitudes (3, nmodes): |a| ~ k exp(-(k/k0)^2) (E(k) ~ k^4 exp(-2 (k/k0)^2))."""
    m = int(np.floor(kmax))
    r = np.arange(-m, m + 1)
    kx, ky, kz = np.meshgrid(np.arange(0, m + 1), r, r, indexing="ij")
    kx, ky, kz = kx.ravel(), ky.ravel(), kz.ravel()
    k2 = kx * kx + ky * ky + kz * kz
    half = (kx > 0) | ((kx == 0) & (ky > 0)) | ((kx == 0) & (ky == 0) & (kz > 0))
    keep = half & (k2 > 0) & (k2 <= kmax * kmax)
    kx, ky, kz, k2 = kx[keep], ky[keep], kz[keep], k2[keep].astype(np.float64)
    k = np.sqrt(k2)
    amp = k * np.exp(-(k / k0) ** 2)
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((3, k.size)) + 1j * rng.standard_normal((3, k.size))) * amp
    kd = (kx * a[0] + ky * a[1] + kz * a[2]) / k2
    a = a - np.stack([kx, ky, kz]) * kd
    return np.stack([kx, ky, kz]), a


def synthesize(kvec, a, shape, zchunk=64):
    """u(z, y, x) = sum over the half-space modes of 2 Re(a e^{2 pi i k.x / n}),
    one component, evaluated separably; returns float64 (nz, ny, nx)."""
    nz, ny, nx = shape
    kx, ky, kz = kvec
    ex = np.exp(2j * np.pi * np.outer(np.arange(kx.min(), kx.max() + 1), np.arange(nx)) / nx)
    eyr = np.arange(ky.min(), ky.max() + 1)
    ey = np.exp(2j * np.pi * np.outer(eyr, np.arange(ny)) / ny)
    kzs = np.unique(kz)
    ez = np.exp(2j * np.pi * np.outer(np.arange(nz), kzs) / nz)      # (nz, nkz)
    # A2[kz](y, x) = sum_{ky} e^{i ky y} sum_{kx} a e^{i kx x}
    A2 = np.zeros((kzs.size, ny, nx), dtype=np.complex128)
    for iz, kzv in enumerate(kzs):
        sel = kz == kzv
        lines = {}
        for kyv, kxv, av in zip(ky[sel], kx[sel], a[sel]):
            lines.setdefault(int(kyv), np.zeros(nx, dtype=np.complex128))
            lines[int(kyv)] += av * ex[kxv - kx.min()]
        for kyv, line in lines.items():
            A2[iz] += np.outer(ey[kyv - eyr[0]], line)
    flat = A2.reshape(kzs.size, ny * nx)
    out = np.empty((nz, ny, nx))
    for z0 in range(0, nz, zchunk):
        z1 = min(nz, z0 + zchunk)
        out[z0:z1] = 2.0 * (ez[z0:z1] @ flat).real.reshape(z1 - z0, ny, nx)
    return out


def solenoidal_field(shape, k0=4.0, kmax=16.0, u_rms=0.03, seed=54321):
    kvec, a = solenoidal_modes(k0, kmax, seed)
    u = [synthesize(kvec, a[c], shape) for c in range(3)]
    rms = np.sqrt(sum(float(np.mean(c * c)) for c in u) / 3.0)
    scale = u_rms / rms if rms > 0 else 0.0
    for c in u:
        c *= scale
    return tuple(u)


KM = 16   # mode table of the device synthesis: 0 <= kx <= KM, |ky|, |kz| <= KM


def dense_modes(kvec, a):
    """The half-space modes as the device synthesis takes them: complex
    (3, kz + KM, ky + KM, kx) table, zero where no mode is drawn."""
    kx, ky, kz = kvec
    if kx.min() < 0 or max(int(kx.max()), int(np.abs(ky).max()), int(np.abs(kz).max())) > KM:
        raise ValueError(f"mode table holds 0 <= kx <= {KM}, |ky|, |kz| <= {KM}")
    t = np.zeros((3, 2 * KM + 1, 2 * KM + 1, KM + 1), dtype=np.complex128)
    t[:, kz + KM, ky + KM, kx] = a
    return t


def parseval_scale(a, u_rms):
    """Scale giving the field sqrt(mean(u.u)/3) = u_rms: each half-space mode
    contributes 2|a|^2 to the mean square of its component, exactly on any grid
    where no two modes (or a mode and a conjugate) alias, i.e. n > 2 kmax."""
    ms = 2.0 * float(np.sum(np.abs(a) ** 2)) / 3.0
    return u_rms / np.sqrt(ms) if ms > 0 else 0.0


def solenoidal_table(shape, k0=4.0, kmax=16.0, u_rms=0.03, seed=54321):
    """(mode table, scale) of solenoidal_field for the device synthesis
    (Lattice.init_velocity): the same field, to rounding, when every edge
    exceeds 2 kmax."""
    if min(shape) <= 2 * int(np.floor(kmax)):
        raise ValueError("the device synthesis needs every edge > 2 kmax (no aliased modes)")
    kvec, a = solenoidal_modes(k0, kmax, seed)
    return dense_modes(kvec, a), parseval_scale(a, u_rms)


def random_spheres(npart, rad, shape, gap=2.0, wall_clear=0.0, walls_y=False, seed=12345,
                   max_tries=10_000_000):
    """npart sphere centres, uniformly random with rejection (random sequential
    addition): surface gaps >= gap under the periodic images, and with walls on
    the y faces each surface at least wall_clear from them.  Centres are in
    (x, y, z) node units, [0, n) per axis as ypglb holds them."""
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    dmin = 2.0 * rad + gap
    cell = dmin
    ncx, ncy, ncz = max(1, int(nx // cell)), max(1, int(ny // cell)), max(1, int(nz // cell))
    grid = {}
    out = np.empty((npart, 3))
    lo_y = rad + wall_clear if walls_y else 0.0
    hi_y = ny - rad - wall_clear if walls_y else float(ny)
    if hi_y <= lo_y:
        raise ValueError("no room between the walls")
    got = tries = 0
    L = np.array([nx, ny, nz], dtype=float)
    while got < npart:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"placed only {got} of {npart} spheres")
        c = np.array([rng.uniform(0, nx), rng.uniform(lo_y, hi_y), rng.uniform(0, nz)])
        ix, iy, iz = int(c[0] / nx * ncx) % ncx, int(c[1] / ny * ncy) % ncy, int(c[2] / nz * ncz) % ncz
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get(((ix + dx) % ncx, (iy + dy) % ncy, (iz + dz) % ncz), ()):
                        d = c - out[j]
                        for k in (0, 1, 2):
                            if k == 1 and walls_y:
                                continue
                            d[k] -= L[k] * np.round(d[k] / L[k])
                        if d @ d < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            out[got] = c
            grid.setdefault((ix, iy, iz), []).append(got)
            got += 1
    return out
