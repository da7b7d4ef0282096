"""Benchmark workloads (the configurations the measurements are quoted on).

``box_case`` builds the periodic forced box with spheres of the configs:
  config 1: 64^3, nu = 0.005, 8 spheres r = 4, density ratio 5, k0 = 4, u_rms 0.02
  config 2: 256^3, nu = 0.0036, 300 spheres r = 8, density ratio 5, k0 = 4, u_rms 0.03
Forcing acts on the |k| <= sqrt(2) shell with the reference's
Ornstein-Uhlenbeck mode process (gaussian.f90:182-372); spheres start from
rest at seeded non-overlapping positions (min surface gap 2, seed 12345).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .params import LBMParams, IFLOWSEED_REF
from .initial import solenoidal_field


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

    def params(self) -> LBMParams:
        return LBMParams(nx=self.n, ny=self.n, nz=self.n, visc=self.visc, npart=self.npart,
                         rad=self.rad, rhopart=self.rhopart, mingap=2.0, forcing=True,
                         xkf=math.sqrt(2.0) + 1e-9)

    def describe(self) -> str:
        return (f"{self.name}: {self.n}^3 periodic D3Q19 MRT, nu={self.visc}, forced |k|<=sqrt2, "
                f"{self.npart} spheres r={self.rad} rho_p/rho_f={self.rhopart}, "
                f"initial E(k)~k^4 exp(-2(k/{self.k0:g})^2) u_rms={self.u_rms}")


CONFIG1 = BoxCase("config1", 64, 0.005, 8, 4.0, 5.0, 4.0, 0.02)
CONFIG2 = BoxCase("config2", 256, 0.0036, 300, 8.0, 5.0, 4.0, 0.03)


def setup_lattice(case: BoxCase, precision: str, field=None):
    """A Lattice holding the case's initial state, spheres placed, ready to
    step from istep = 1 (particles already released)."""
    from .lbm import Lattice
    prm = case.params()
    lat = Lattice(prm, precision)
    if field is None:
        field = solenoidal_field((prm.nz, prm.ny, prm.nx), k0=case.k0, u_rms=case.u_rms,
                                 seed=IFLOWSEED_REF)
    ux, uy, uz = field
    lat.upload_macro(None, ux, uy, uz)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    if case.npart > 0:
        lat.prt_init(case.npart)
        lat.initpartpos(prm.iseedp, mingap_check=True)
        lat.initpartomg()                       # spheres start from rest
        lat.prt_set("wp", np.zeros((case.npart, 3)))
        lat.beads_links()
    lat.sync()
    return lat, prm, field
