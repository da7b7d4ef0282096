"""ctypes binding of the C oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("ipart", C.c_int), ("npart", C.c_int), ("irelease", C.c_int),
        ("bb_mode", C.c_int), ("forcing_on", C.c_int), ("walls_y", C.c_int),
        ("rad", C.c_double), ("rhopart", C.c_double), ("mingap", C.c_double), ("rho0", C.c_double),
        ("s1", C.c_double), ("s2", C.c_double), ("s4", C.c_double), ("s9", C.c_double),
        ("s10", C.c_double), ("s13", C.c_double), ("s16", C.c_double),
        ("omegepsl", C.c_double), ("omegepslj", C.c_double), ("omegxx", C.c_double),
        ("ww0", C.c_double), ("ww1", C.c_double), ("ww2", C.c_double),
        ("volp", C.c_double), ("amp", C.c_double), ("aip", C.c_double), ("rhog", C.c_double),
        ("stf0", C.c_double), ("stf1", C.c_double), ("gscale", C.c_double),
        ("vscale", C.c_double), ("tscale", C.c_double), ("xkf", C.c_double), ("pi2", C.c_double),
        ("body_force", C.c_double * 3), ("particle_force", C.c_double * 3),
        ("mingap_w", C.c_double), ("stf0_w", C.c_double), ("stf1_w", C.c_double),
    ]


_PTR_FIELDS_D = ["f", "rho", "ux", "uy", "uz", "fx", "fy", "fz"]
_PTR_FIELDS_I = ["ibnodes", "isnodes", "ibnodes0", "isnodes0", "ilink", "mlink"]
_PART_FIELDS = ["yp", "ypglb", "ypglbp", "wp", "wpp", "omgp", "omgpp", "fHIp", "flubp", "torqp",
                "dwdt", "domgdt", "forcep", "forcepp", "torqpp", "thetap"]


class OrcState(C.Structure):
    _fields_ = (
        [("f", _D), ("rho", _D), ("ux", _D), ("uy", _D), ("uz", _D),
         ("fx", _D), ("fy", _D), ("fz", _D),
         ("ibnodes", _I), ("isnodes", _I), ("ibnodes0", _I), ("isnodes0", _I),
         ("ilink", _I), ("mlink", _I), ("alink", _D)]
        + [(n, _D) for n in _PART_FIELDS]
        + [("a1r", C.c_double * 150), ("a2r", C.c_double * 150), ("a3r", C.c_double * 150),
           ("b1r", C.c_double * 150), ("b2r", C.c_double * 150), ("b3r", C.c_double * 150),
           ("ivf", C.c_int * 32), ("ivp", C.c_int * 32),
           ("iseedf", C.c_int), ("iyf", C.c_int), ("iseedp", C.c_int), ("iyp", C.c_int),
           ("istep", C.c_int), ("istep00", C.c_int),
           ("nconflict", C.c_int), ("alpha_faults", C.c_int), ("novlp", C.c_int)]
    )


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle/build.sh)")
        _lib = C.CDLL(LIB_PATH)
        P, S = C.POINTER(OrcParams), C.POINTER(OrcState)
        for name in ["orc_forcing_modes", "orc_forcing", "orc_initpop", "orc_collision_mrt",
                     "orc_streaming", "orc_rhoupdat", "orc_macrovar", "orc_initpartvel",
                     "orc_initpartomg", "orc_beads_links", "orc_beads_collision",
                     "orc_beads_lubforce", "orc_beads_move", "orc_beads_filling",
                     "orc_release"]:
            fn = getattr(_lib, name)
            fn.argtypes = [P, S]
            fn.restype = None
        for name in ["orc_initpartpos", "orc_beads_redistribute", "orc_timestep"]:
            fn = getattr(_lib, name)
            fn.argtypes = [P, S]
            fn.restype = C.c_int
        _lib.orc_ovlpchk.argtypes = [P, S, _I]
        _lib.orc_ovlpchk.restype = C.c_int
        _lib.orc_ranf.argtypes = [_I, _I, _I]
        _lib.orc_ranf.restype = C.c_double
        _lib.orc_force_field.argtypes = [P, _D, _D, _D, _D, _D, _D]
        _lib.orc_force_field.restype = None
        _lib.orc_feqpnt.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double, _D]
        _lib.orc_feqpnt.restype = None
        _lib.orc_nthreads.restype = C.c_int
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class Oracle:
    """Owns numpy arrays in the reference layout and calls the C oracle.

    Arrays: ``f`` has shape (nz, ny, nx, 19) (== Fortran f(0:18,lx,ly,lz));
    scalar fields (nz, ny, nx); link arrays (nz, ny, nx, 18); particle arrays
    (npart, 3)."""

    def __init__(self, prm, bb_mode: int = 1, ipart=None):
        self.prm = prm
        nx, ny, nz = prm.nx, prm.ny, prm.nz
        npart = prm.npart
        self.P = OrcParams()
        for name, _ in OrcParams._fields_:
            if name in ("body_force", "particle_force"):
                v = getattr(prm, name, (0.0, 0.0, 0.0))
                for k in range(3):
                    getattr(self.P, name)[k] = float(v[k])
            elif hasattr(prm, name):
                setattr(self.P, name, getattr(prm, name))
        self.P.ipart = int(npart > 0 if ipart is None else ipart)
        self.P.bb_mode = bb_mode
        self.P.forcing_on = int(prm.forcing)
        self.P.walls_y = int(prm.walls_y)
        shp = (nz, ny, nx)
        self.f = np.zeros(shp + (19,))
        for n in ["rho", "ux", "uy", "uz", "fx", "fy", "fz"]:
            setattr(self, n, np.zeros(shp))
        for n in ["ibnodes", "isnodes", "ibnodes0", "isnodes0"]:
            setattr(self, n, np.full(shp, -1, dtype=np.int32))
        self.ilink = np.full(shp + (18,), -1, dtype=np.int32)
        self.mlink = np.full(shp + (18,), -1, dtype=np.int32)
        self.alink = np.zeros(shp + (18,))
        np_ = max(npart, 1)
        for n in _PART_FIELDS:
            setattr(self, n, np.zeros((np_, 3)))
        self.S = OrcState()
        self.S.iseedf = -int(prm.iseedf)   # para.f90:281 negates the seed
        self.S.iseedp = int(prm.iseedp)
        self.S.istep = 0
        self.S.istep00 = 1
        self._bind()

    def _bind(self):
        S = self.S
        for n in _PTR_FIELDS_D:
            setattr(S, n, _dp(getattr(self, n)))
        for n in _PTR_FIELDS_I:
            setattr(S, n, _ip(getattr(self, n)))
        S.alink = _dp(self.alink)
        for n in _PART_FIELDS:
            setattr(S, n, _dp(getattr(self, n)))

    def call(self, name):
        return getattr(lib(), "orc_" + name)(C.byref(self.P), C.byref(self.S))

    # convenience wrappers mirroring the reference subroutine names
    def forcing(self): self.call("forcing")
    def initpop(self): self.call("initpop")
    def collision_mrt(self): self.call("collision_mrt")
    def streaming(self): self.call("streaming")
    def rhoupdat(self): self.call("rhoupdat")
    def macrovar(self): self.call("macrovar")
    def beads_links(self): self.call("beads_links")
    def beads_collision(self): self.call("beads_collision")
    def beads_lubforce(self): self.call("beads_lubforce")
    def beads_move(self): self.call("beads_move")
    def beads_redistribute(self): return self.call("beads_redistribute")
    def beads_filling(self): self.call("beads_filling")
    def initpartpos(self): return self.call("initpartpos")
    def initpartvel(self): self.call("initpartvel")
    def release(self): self.call("release")

    def timestep(self):
        """One pass of the time-loop body for the next step number."""
        self.S.istep += 1
        return self.call("timestep")

    def modes(self):
        return (np.ctypeslib.as_array(self.S.a1r).copy(), np.ctypeslib.as_array(self.S.a2r).copy(),
                np.ctypeslib.as_array(self.S.a3r).copy())

    def set_modes(self, a1r, a2r, a3r):
        for dst, src in ((self.S.a1r, a1r), (self.S.a2r, a2r), (self.S.a3r, a3r)):
            np.ctypeslib.as_array(dst)[:] = np.asarray(src, dtype=np.float64).ravel()

    def force_field(self, a1r, a2r, a3r):
        a = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (a1r, a2r, a3r)]
        lib().orc_force_field(C.byref(self.P), _dp(a[0]), _dp(a[1]), _dp(a[2]),
                              _dp(self.fx), _dp(self.fy), _dp(self.fz))


def ranf_sequence(seed: int, n: int):
    """n draws of the reference RNG started from ``seed`` (negative = reseed)."""
    idum = C.c_int(seed)
    iy = C.c_int(0)
    iv = (C.c_int * 32)()
    out = np.empty(n)
    for k in range(n):
        out[k] = lib().orc_ranf(C.byref(idum), C.byref(iy), iv)
    return out
