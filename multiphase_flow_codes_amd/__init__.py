"""MI355X-native D3Q19 MRT lattice-Boltzmann step with finite-size moving
spheres: a drop-in for the forced-turbulence particle code's hot path
(collision_MRT, streaming, macrovar, FORCING, beads_*) behind a C ABI."""
from .params import LBMParams, CIX, CIY, CIZ, IPOPP  # noqa: F401

__all__ = ["LBMParams", "CIX", "CIY", "CIZ", "IPOPP"]
