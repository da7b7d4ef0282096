"""Physical and lattice constants of the D3Q19 MRT step.

Host-side mirror of the reference's ``para`` subroutine
(3DBox_Forced_HI_Turbulence_Moving_Particles_LBM/para.f90:1-288): given the
lattice size, the viscosity and the particle/forcing settings it derives every
constant the step needs (relaxation time, MRT rates and equilibrium weights,
lattice weights, particle volume/mass/inertia, lubrication constants, forcing
time scales) with the same formulas and in the same floating-point order, so
the library receives exactly the numbers the Fortran driver would hold in
``module var_inc``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, asdict

# D3Q19 velocity set and opposite-direction table (para.f90:171-175)
CIX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
CIY = (0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1)
CIZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1)
IPOPP = (0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15)

# reference defaults (para.f90:14-37, 237-247; var_inc.f90:19-28)
VSCALE_REF = 2.037183272e-3
ISEEDF_REF = 232300
ISEEDP_REF = 12345
IFLOWSEED_REF = 54321


@dataclass
class LBMParams:
    """Run constants.  Derived fields are filled by :meth:`derive`."""

    nx: int
    ny: int
    nz: int
    visc: float                  # lattice kinematic viscosity
    mrt_type: int = 1            # para.f90:96-135
    rho0: float = 1.0
    # particles
    npart: int = 0
    rad: float = 4.0
    rhopart: float = 5.0         # particle density (rho0 = 1)
    mingap: float = 2.0
    gscale: float = 10000.0 / 4.0
    rhog: float = 0.0            # buoyancy-corrected gravity (para.f90:239)
    stf0: float = 0.025
    stf1: float = 0.002
    irelease: int = 1
    iseedp: int = ISEEDP_REF
    # forcing
    forcing: bool = True
    vscale: float = VSCALE_REF
    xkf: float = math.sqrt(8.0)  # |k| cut-off of the forced shell (gaussian.f90:195)
    iseedf: int = ISEEDF_REF
    # uniform body force added to the forcing field (fluid and particles)
    body_force: tuple = (0.0, 0.0, 0.0)
    # uniform force per unit sphere volume on the spheres only
    particle_force: tuple = (0.0, 0.0, 0.0)
    # half-way bounce-back walls bounding the y direction
    walls_y: bool = False
    # sphere-wall lubrication range (also the spheres' placement clearance from
    # the walls) and stiffness (3DChannel_.../var_inc.f90:31, para.f90:241-242)
    mingap_w: float = 2.0
    stf0_w: float = 0.025
    stf1_w: float = 0.002
    # derived (para.f90)
    tau: float = field(default=0.0, init=False)
    pi2: float = field(default=0.0, init=False)
    tscale: float = field(default=0.0, init=False)
    s1: float = field(default=0.0, init=False)
    s2: float = field(default=0.0, init=False)
    s4: float = field(default=0.0, init=False)
    s9: float = field(default=0.0, init=False)
    s10: float = field(default=0.0, init=False)
    s13: float = field(default=0.0, init=False)
    s16: float = field(default=0.0, init=False)
    omegepsl: float = field(default=0.0, init=False)
    omegepslj: float = field(default=0.0, init=False)
    omegxx: float = field(default=0.0, init=False)
    ww0: float = field(default=0.0, init=False)
    ww1: float = field(default=0.0, init=False)
    ww2: float = field(default=0.0, init=False)
    volp: float = field(default=0.0, init=False)
    amp: float = field(default=0.0, init=False)
    aip: float = field(default=0.0, init=False)

    def __post_init__(self):
        self.derive()

    @classmethod
    def from_reference_scaling(cls, n: int, anu: float = 0.0945, vscale: float = VSCALE_REF, **kw):
        """Viscosity from the spectral-code scaling visc = anu*vscale*N/(2 pi)
        (para.f90:91)."""
        pi2 = 2.0 * (4.0 * math.atan(1.0))
        visc = anu * vscale * float(n) / pi2
        return cls(nx=n, ny=n, nz=n, visc=visc, vscale=vscale, **kw)

    def derive(self) -> "LBMParams":
        pi = 4.0 * math.atan(1.0)
        self.pi2 = 2.0 * pi
        self.tscale = self.pi2 * self.vscale / float(self.ny)
        self.tau = 3.0 * self.visc + 0.5
        if self.mrt_type == 1:
            self.omegepsl, self.omegepslj, self.omegxx = 0.0, -475.0 / 63.0, 0.0
        elif self.mrt_type in (2, 3):
            self.omegepsl, self.omegepslj, self.omegxx = 3.0, -11.0 / 2.0, -1.0 / 2.0
        else:
            raise ValueError(f"mrt_type must be 1, 2 or 3, got {self.mrt_type}")
        self.s9 = 1.0 / self.tau
        self.s13 = self.s9
        if self.mrt_type == 1:
            self.s1, self.s2, self.s4, self.s10, self.s16 = 1.19, 1.4, 1.2, 1.4, 1.98
        elif self.mrt_type == 2:
            self.s1 = self.s2 = self.s4 = self.s10 = self.s16 = self.s9
        else:
            self.s1 = 1.8
            self.s2 = self.s1
            self.s4 = self.s9
            self.s10 = self.s1
            self.s16 = self.s1
        self.ww0, self.ww1, self.ww2 = 1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0
        r = self.rad
        self.volp = 4.0 / 3.0 * pi * (r * r * r)
        self.amp = self.rhopart * self.volp
        self.aip = 0.4 * self.amp * (r * r)
        return self

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def as_dict(self) -> dict:
        return asdict(self)
