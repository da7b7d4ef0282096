"""Benchmark workloads (the cases the measurements are quoted on).

``BoxCase`` builds the boxes with spheres:
  box64:       64^3 periodic, nu = 0.005, 8 spheres r = 4, density ratio 5,
               k0 = 4, u_rms 0.02, forced
  box256:      256^3 periodic, nu = 0.0036, 300 spheres r = 8, density ratio 5,
               k0 = 4, u_rms 0.03, forced
  channel1024: 1024^3 channel (half-way bounce-back walls on the y faces),
               decaying (no forcing), nu = 0.002, 20,000 spheres r = 6,
               density ratio 2.5, k0 = 8, u_rms 0.03, sphere surfaces at least
               8 nodes from the walls
  channel256:  a 256^3 section of channel1024 with its solid fraction (312
               spheres), for samples of that case on the host
Forcing acts on the |k| <= sqrt(2) shell with the reference's
Ornstein-Uhlenbeck mode process (gaussian.f90:182-372).  Spheres start from
rest at seeded non-overlapping positions (min surface gap 2): the reference's
initpartpos (seed 12345) in the periodic boxes, random sequential addition
with the wall clearance in the channel (the reference's channel placement
keeps only rad + mingap_w = rad + 2 from the walls, CHAN/partlib.f90:10-24).
The initial velocity is the solenoidal field of initial.solenoidal_modes
(seed 54321), synthesised on the device.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .params import LBMParams, IFLOWSEED_REF
from .initial import random_spheres, solenoidal_field, solenoidal_table


@dataclass
class BoxCase:
    name: str
    n: int
    visc: float
    npart: int
    rad: float
    rhopart: float
    k0: float
    u_rms: float
    walls: bool = False           # half-way bounce-back walls on the y faces, no forcing
    wall_clear: float = 0.0       # placement: sphere surface to wall distance (walls only)

    def params(self) -> LBMParams:
        if self.walls:
            return LBMParams(nx=self.n, ny=self.n, nz=self.n, visc=self.visc, npart=self.npart,
                             rad=self.rad, rhopart=self.rhopart, mingap=2.0, forcing=False,
                             walls_y=True, mingap_w=2.0)
        return LBMParams(nx=self.n, ny=self.n, nz=self.n, visc=self.visc, npart=self.npart,
                         rad=self.rad, rhopart=self.rhopart, mingap=2.0, forcing=True,
                         xkf=math.sqrt(2.0) + 1e-9)

    @property
    def shape(self):
        return (self.n, self.n, self.n)

    def describe(self) -> str:
        if self.walls:
            return (f"{self.name}: {self.n}^3 channel D3Q19 MRT (half-way bounce-back walls at "
                    f"y=0 and y={self.n - 1}), decaying, nu={self.visc}, {self.npart} spheres "
                    f"r={self.rad} rho_p/rho_f={self.rhopart} >= {self.wall_clear:g} nodes from the "
                    f"walls, initial E(k)~k^4 exp(-2(k/{self.k0:g})^2) u_rms={self.u_rms}")
        return (f"{self.name}: {self.n}^3 periodic D3Q19 MRT, nu={self.visc}, forced |k|<=sqrt2, "
                f"{self.npart} spheres r={self.rad} rho_p/rho_f={self.rhopart}, "
                f"initial E(k)~k^4 exp(-2(k/{self.k0:g})^2) u_rms={self.u_rms}")

    def host_field(self):
        """The initial velocity on the host (for the CPU oracle)."""
        return solenoidal_field(self.shape, k0=self.k0, u_rms=self.u_rms, seed=IFLOWSEED_REF)


BOX64 = BoxCase("box64", 64, 0.005, 8, 4.0, 5.0, 4.0, 0.02)
BOX256 = BoxCase("box256", 256, 0.0036, 300, 8.0, 5.0, 4.0, 0.03)
CHANNEL1024 = BoxCase("channel1024", 1024, 0.002, 20000, 6.0, 2.5, 8.0, 0.03, walls=True, wall_clear=8.0)
WORKLOADS = {c.name: c for c in (BOX64, BOX256, CHANNEL1024)}


# a 256^3 section of the channel: the same walls, viscosity, sphere radius,
# density ratio and solid fraction (312 spheres), for host-side samples
CHANNEL256 = BoxCase("channel256", 256, 0.002, 312, 6.0, 2.5, 8.0, 0.03, walls=True, wall_clear=8.0)


def setup_lattice(case: BoxCase, precision: str, field=None, storage: str = "ab",
                  perturb: float = 0.0):
    """A Lattice holding the case's initial state, spheres placed, ready to
    step from istep = 1 (particles already released).  `field` (ux, uy, uz
    on the host) replaces the device synthesis of the initial velocity.
    `perturb` > 0 multiplies every mode amplitude of the synthesised field by
    (1 + perturb * xi), xi standard normal (seed 7): a control run whose
    initial state differs from the case's at that relative level."""
    from .lbm import Lattice
    prm = case.params()
    lat = Lattice(prm, precision, storage=storage)
    if field is None:
        table, scale = solenoidal_table(case.shape, k0=case.k0, u_rms=case.u_rms, seed=IFLOWSEED_REF)
        if perturb > 0.0:
            xi = np.random.default_rng(7).standard_normal(table.shape)
            table = table * (1.0 + perturb * xi)
        lat.init_velocity(table, scale)
    else:
        ux, uy, uz = field
        lat.upload_macro(None, ux, uy, uz)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    if case.npart > 0:
        lat.prt_init(case.npart)
        if case.walls:
            pos = random_spheres(case.npart, case.rad, case.shape, gap=prm.mingap,
                                 wall_clear=case.wall_clear, walls_y=True, seed=prm.iseedp)
            lat.prt_set("ypglb", pos)
        else:
            lat.initpartpos(prm.iseedp, mingap_check=True)
        lat.initpartomg()                       # spheres start from rest
        lat.prt_set("wp", np.zeros((case.npart, 3)))
        lat.beads_links()
    lat.sync()
    return lat, prm, field
