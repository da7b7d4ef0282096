"""ctypes binding of libmfc_lbm (the HIP library) with methods named after the
reference subroutines they replace (BOX/ = 3DBox_Forced_HI_Turbulence_Moving_
Particles_LBM/).  Arrays use the reference layout reshaped C-order:
``f`` (nz, ny, nx, 19) == Fortran f(0:18,lx,ly,lz); fields (nz, ny, nx);
particle arrays (npart, 3) == Fortran (3, npart).

The library is loaded from the package directory; if it is missing this
module raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import LBMParams

HERE = os.path.dirname(os.path.abspath(__file__))
# MFC_LIB: another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("MFC_LIB") or os.path.join(HERE, "libmfc_lbm.so")

MFC_OK = 0
ERRORS = {1: "MFC_ERR_ARG", 2: "MFC_ERR_HIP", 3: "MFC_ERR_STATE", 4: "MFC_ERR_OVERLAP",
          5: "MFC_ERR_CAPACITY", 6: "MFC_ERR_NONFINITE", 7: "MFC_ERR_NODEVICE"}
PREC = {"fp64": 0, "fp32": 1, "fp32x": 2}   # fp32x: FP32 storage, FP64 collision
STORAGE = {"ab": 0, "aa": 1}                # two buffers, or one updated in place (AA pattern)
PFIELDS = {"ypglb": 0, "ypglbp": 1, "wp": 2, "wpp": 3, "omgp": 4, "omgpp": 5, "fHIp": 6,
           "flubp": 7, "torqp": 8, "forcepp": 9, "torqpp": 10, "dwdt": 11, "domgdt": 12,
           "thetap": 13}
STATUS_NAMES = ["overlap", "overlap_step", "link_overflow", "alpha_fault", "contact_links",
                "nonfinite", "refill", "nlinks", "nsolid", "box_fault", "cell_overflow", "fast",
                "order_cap"]


class MfcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class MfcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("precision", C.c_int32), ("npart_max", C.c_int32), ("walls_y", C.c_int32),
        ("iseedf", C.c_int32), ("irelease", C.c_int32), ("storage", C.c_int32),
        ("s1", C.c_double), ("s2", C.c_double), ("s4", C.c_double), ("s9", C.c_double),
        ("s10", C.c_double), ("s13", C.c_double), ("s16", C.c_double),
        ("omegepsl", C.c_double), ("omegepslj", C.c_double), ("omegxx", C.c_double),
        ("ww0", C.c_double), ("ww1", C.c_double), ("ww2", C.c_double), ("rho0", C.c_double),
        ("rad", C.c_double), ("rhopart", C.c_double), ("mingap", C.c_double),
        ("volp", C.c_double), ("amp", C.c_double), ("aip", C.c_double), ("rhog", C.c_double),
        ("stf0", C.c_double), ("stf1", C.c_double), ("gscale", C.c_double),
        ("vscale", C.c_double), ("tscale", C.c_double), ("xkf", C.c_double), ("pi2", C.c_double),
        ("body_force", C.c_double * 3),
        ("particle_force", C.c_double * 3),
        ("mingap_w", C.c_double), ("stf0_w", C.c_double), ("stf1_w", C.c_double),
    ]


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_lib = None

# (name, argtypes) of every exported entry point; include/mfc_lbm.h is the source of truth
SIGNATURES = {
    "mfc_create": [C.POINTER(MfcParams), C.POINTER(C.c_void_p)],
    "mfc_destroy": [C.c_void_p],
    "mfc_last_error": [C.c_void_p],
    "mfc_sync": [C.c_void_p],
    "mfc_status": [C.c_void_p, _I],
    "mfc_set_stream": [C.c_void_p, C.c_void_p],
    "mfc_get_stream": [C.c_void_p],
    "mfc_upload_f": [C.c_void_p, _D],
    "mfc_download_f": [C.c_void_p, _D],
    "mfc_upload_macro": [C.c_void_p, _D, _D, _D, _D],
    "mfc_download_macro": [C.c_void_p, _D, _D, _D, _D],
    "mfc_download_force": [C.c_void_p, _D, _D, _D],
    "mfc_get_forcing_state": [C.c_void_p, _I, _I, _I, _D, _D, _D, _D, _D, _D],
    "mfc_set_forcing_state": [C.c_void_p, C.c_int32, C.c_int32, _I, _D, _D, _D, _D, _D, _D],
    "mfc_initpop": [C.c_void_p],
    "mfc_init_velocity": [C.c_void_p, _D, C.c_double],
    "mfc_forcing": [C.c_void_p],
    "mfc_collision_mrt": [C.c_void_p],
    "mfc_streaming": [C.c_void_p],
    "mfc_rhoupdat": [C.c_void_p, _D],
    "mfc_macrovar": [C.c_void_p],
    "mfc_prt_init": [C.c_void_p, C.c_int32],
    "mfc_prt_set": [C.c_void_p, C.c_int32, _D],
    "mfc_prt_get": [C.c_void_p, C.c_int32, _D],
    "mfc_initpartpos": [C.c_void_p, _I, C.c_int32],
    "mfc_initpartvel": [C.c_void_p],
    "mfc_initpartomg": [C.c_void_p],
    "mfc_beads_links": [C.c_void_p],
    "mfc_beads_collision": [C.c_void_p],
    "mfc_set_bb_order": [C.c_void_p, C.c_int32],
    "mfc_beads_lubforce": [C.c_void_p],
    "mfc_beads_move": [C.c_void_p, C.c_int32],
    "mfc_beads_redistribute": [C.c_void_p, C.c_int32],
    "mfc_beads_filling": [C.c_void_p],
    "mfc_get_maps": [C.c_void_p, C.c_int32, _I, _I],
    "mfc_get_links": [C.c_void_p, _I, _I, _D],
    "mfc_step": [C.c_void_p, C.c_int32, C.c_int32],
    "mfc_diag_sums": [C.c_void_p, _D],
    "mfc_diag_sums_noalloc": [C.c_void_p, C.c_void_p],
    "mfc_diag_spectrum": [C.c_void_p, C.c_int32, _D],
    "mfc_set_timing": [C.c_void_p, C.c_int32],
    "mfc_get_timing": [C.c_void_p, _D, C.POINTER(C.c_int64)],
    "mfc_ens_create": [C.POINTER(C.c_void_p)],
    "mfc_ens_destroy": [C.c_void_p],
    "mfc_ens_last_error": [C.c_void_p],
    "mfc_ens_add": [C.c_void_p, C.c_void_p, C.c_int32],
    "mfc_ens_remove": [C.c_void_p, C.c_void_p],
    "mfc_ens_size": [C.c_void_p],
    "mfc_ens_step": [C.c_void_p, C.c_int32],
    "mfc_ens_sync": [C.c_void_p],
    "mfc_ens_set_block": [C.c_void_p, C.c_int32],
    "mfc_ens_set_graph": [C.c_void_p, C.c_int32],
    "mfc_ens_set_tile_order": [C.c_void_p, C.c_int32],
    "mfc_ens_autotune": [C.c_void_p, C.c_int32, _I, _D],
    "mfc_ens_host_time": [C.c_void_p, _D, C.POINTER(C.c_int64)],
    "mfc_ens_get_stream": [C.c_void_p],
}


def lib():
    """Load libmfc_lbm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmfc_lbm.so not found at {LIB_PATH}; run "
                               "`python -m multiphase_flow_codes_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, argt in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = C.c_int
        L.mfc_last_error.restype = C.c_char_p
        L.mfc_get_stream.restype = C.c_void_p
        L.mfc_ens_last_error.restype = C.c_char_p
        L.mfc_ens_get_stream.restype = C.c_void_p
        _lib = L
    return _lib


def make_params(prm: LBMParams, precision: str = "fp64", npart_max: int | None = None,
                storage: str = "ab") -> MfcParams:
    p = MfcParams()
    p.storage = STORAGE[storage]
    p.nx, p.ny, p.nz = prm.nx, prm.ny, prm.nz
    p.precision = PREC[precision]
    p.npart_max = prm.npart if npart_max is None else npart_max
    p.walls_y = int(prm.walls_y)
    p.iseedf = -int(prm.iseedf)   # para.f90:281 negates the seed
    p.irelease = prm.irelease
    for name in ["s1", "s2", "s4", "s9", "s10", "s13", "s16", "omegepsl", "omegepslj", "omegxx",
                 "ww0", "ww1", "ww2", "rho0", "rad", "rhopart", "mingap", "volp", "amp", "aip",
                 "rhog", "stf0", "stf1", "gscale", "vscale", "tscale", "pi2", "mingap_w", "stf0_w",
                 "stf1_w"]:
        setattr(p, name, float(getattr(prm, name)))
    p.xkf = float(prm.xkf) if prm.forcing else 0.0
    for k in range(3):
        p.body_force[k] = float(prm.body_force[k])
        p.particle_force[k] = float(prm.particle_force[k])
    return p


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class Lattice:
    """One box on the GPU (a context of libmfc_lbm)."""

    def __init__(self, prm: LBMParams, precision: str = "fp64", storage: str = "ab"):
        self.prm = prm
        self.precision = precision
        self.storage = storage
        self.shape = (prm.nz, prm.ny, prm.nx)
        self._p = make_params(prm, precision, storage=storage)
        h = C.c_void_p()
        rc = lib().mfc_create(C.byref(self._p), C.byref(h))
        if rc != MFC_OK:
            raise MfcError(rc, "mfc_create failed")
        self.h = h
        self.npart = 0

    # -- plumbing ------------------------------------------------------------
    def _chk(self, rc):
        if rc != MFC_OK:
            msg = lib().mfc_last_error(self.h)
            raise MfcError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            lib().mfc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._chk(lib().mfc_sync(self.h))

    def status(self) -> dict:
        out = np.zeros(16, dtype=np.int32)
        self._chk(lib().mfc_status(self.h, _ip(out)))
        return {n: int(out[i]) for i, n in enumerate(STATUS_NAMES)}

    def stream_handle(self) -> int:
        return lib().mfc_get_stream(self.h)

    # -- transfers -----------------------------------------------------------
    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.shape + (19,)
        self._chk(lib().mfc_upload_f(self.h, _dp(f)))

    def download_f(self):
        f = np.empty(self.shape + (19,))
        self._chk(lib().mfc_download_f(self.h, _dp(f)))
        return f

    def upload_macro(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, uz)]
        ptrs = [None if a is None else _dp(a) for a in arrs]
        self._chk(lib().mfc_upload_macro(self.h, *ptrs))

    def download_macro(self):
        out = [np.empty(self.shape) for _ in range(4)]
        self._chk(lib().mfc_download_macro(self.h, *[_dp(a) for a in out]))
        return tuple(out)

    def download_force(self):
        out = [np.empty(self.shape) for _ in range(3)]
        self._chk(lib().mfc_download_force(self.h, *[_dp(a) for a in out]))
        return tuple(out)

    def get_forcing_state(self):
        i1, i2 = C.c_int32(), C.c_int32()
        iv = np.zeros(32, dtype=np.int32)
        arrs = [np.zeros(150) for _ in range(6)]
        self._chk(lib().mfc_get_forcing_state(self.h, C.byref(i1), C.byref(i2), _ip(iv),
                                              *[_dp(a) for a in arrs]))
        return {"iseedf": i1.value, "iyf": i2.value, "ivf": iv, "b": arrs[:3], "a": arrs[3:]}

    def set_forcing_state(self, iseedf, iyf, ivf, b, a=None):
        iv = np.ascontiguousarray(ivf, dtype=np.int32)
        bs = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in b]
        as_ = [None] * 3 if a is None else [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in a]
        self._chk(lib().mfc_set_forcing_state(self.h, int(iseedf), int(iyf), _ip(iv),
                                              *[_dp(x) for x in bs],
                                              *[None if x is None else _dp(x) for x in as_]))

    # -- the reference subroutines ---------------------------------------------
    def initpop(self): self._chk(lib().mfc_initpop(self.h))

    def init_velocity(self, amp, scale: float):
        """Band-limited velocity field synthesised on the device from a dense
        mode table (see initial.dense_modes), rho = 0; then initpop."""
        a = np.ascontiguousarray(amp, dtype=np.complex128)
        assert a.shape == (3, 33, 33, 17), a.shape
        self._chk(lib().mfc_init_velocity(self.h, a.ctypes.data_as(_D), float(scale)))
    def forcing(self): self._chk(lib().mfc_forcing(self.h))
    def collision_mrt(self): self._chk(lib().mfc_collision_mrt(self.h))
    def streaming(self): self._chk(lib().mfc_streaming(self.h))
    def macrovar(self): self._chk(lib().mfc_macrovar(self.h))

    def rhoupdat(self) -> float:
        d = C.c_double()
        self._chk(lib().mfc_rhoupdat(self.h, C.byref(d)))
        return d.value

    def prerelax(self, rhoepsl: float, max_iter: int = 15000) -> tuple:
        """Pre-relaxation of the density at frozen velocity (BOX/main.f90:86-108):
        collision_MRT, streaming, rhoupdat until max|rho - rho_prev| <= rhoepsl or
        the iteration count passes max_iter; then macrovar.  Returns
        (iterations, final residual)."""
        it = 0
        while True:
            self.collision_mrt()
            self.streaming()
            err = self.rhoupdat()
            if err <= rhoepsl or it > max_iter:
                break
            it += 1
        self.macrovar()
        return it, err

    # -- particles ---------------------------------------------------------------
    def prt_init(self, npart: int):
        self._chk(lib().mfc_prt_init(self.h, int(npart)))
        self.npart = int(npart)

    def prt_set(self, field: str, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.npart, 3)
        self._chk(lib().mfc_prt_set(self.h, PFIELDS[field], _dp(a)))

    def prt_get(self, field: str):
        a = np.empty((self.npart, 3))
        self._chk(lib().mfc_prt_get(self.h, PFIELDS[field], _dp(a)))
        return a

    def initpartpos(self, iseedp: int, mingap_check: bool = True) -> int:
        s = C.c_int32(int(iseedp))
        self._chk(lib().mfc_initpartpos(self.h, C.byref(s), int(mingap_check)))
        return s.value

    def initpartvel(self): self._chk(lib().mfc_initpartvel(self.h))
    def initpartomg(self): self._chk(lib().mfc_initpartomg(self.h))
    def beads_links(self): self._chk(lib().mfc_beads_links(self.h))
    def beads_collision(self): self._chk(lib().mfc_beads_collision(self.h))
    def beads_lubforce(self): self._chk(lib().mfc_beads_lubforce(self.h))
    def beads_move(self, istep: int): self._chk(lib().mfc_beads_move(self.h, int(istep)))
    def beads_redistribute(self, istep: int): self._chk(lib().mfc_beads_redistribute(self.h, int(istep)))
    def beads_filling(self): self._chk(lib().mfc_beads_filling(self.h))

    def get_maps(self, which: int = 0):
        ib = np.empty(self.shape, dtype=np.int32)
        isn = np.empty(self.shape, dtype=np.int32)
        self._chk(lib().mfc_get_maps(self.h, int(which), _ip(ib), _ip(isn)))
        return ib, isn

    def get_links(self):
        il = np.empty(self.shape + (18,), dtype=np.int32)
        ml = np.empty(self.shape + (18,), dtype=np.int32)
        al = np.empty(self.shape + (18,))
        self._chk(lib().mfc_get_links(self.h, _ip(il), _ip(ml), _dp(al)))
        return il, ml, al

    # -- whole step ------------------------------------------------------------------
    def release(self, iseedp: int | None = None):
        """Particle release (BOX/main.f90:149-165): positions, velocities,
        zero spin, node maps."""
        seed = self.prm.iseedp if iseedp is None else iseedp
        self.initpartpos(seed, mingap_check=True)
        self.initpartvel()
        self.initpartomg()
        self.beads_links()

    def step(self, istep: int, nsteps: int = 1):
        self._chk(lib().mfc_step(self.h, int(istep), int(nsteps)))

    def diag_sums(self) -> dict:
        """FP64 sums over all nodes of the macrovar fields (device reduction)."""
        out = np.zeros(6)
        self._chk(lib().mfc_diag_sums(self.h, _dp(out)))
        return {"rho": out[0], "ux": out[1], "uy": out[2], "uz": out[3], "u2": out[4],
                "u2max": out[5]}

    def diag_sums_noalloc(self, dev_out6: int):
        """mfc_diag_sums into six doubles of device memory at address
        dev_out6, with the library's static scratch (allocates nothing,
        does not synchronise)."""
        self._chk(lib().mfc_diag_sums_noalloc(self.h, C.c_void_p(int(dev_out6))))

    def diag_spectrum(self, nek: int) -> np.ndarray:
        """Kinetic-energy spectrum E(ik), ik = 1..nek, computed on the device."""
        ek = np.zeros(int(nek))
        self._chk(lib().mfc_diag_spectrum(self.h, int(nek), _dp(ek)))
        return ek

    def set_bb_order(self, order: str):
        """Bounce-back link order: "snapshot" (every link against one
        post-collision field, the default) or "reference" (the reference's
        one-rank loop order, partlib.f90:1363-1621, which differs only for
        links whose own node is solid for another sphere)."""
        codes = {"snapshot": 0, "reference": 1}
        if order not in codes:
            raise ValueError(f"bounce-back order must be one of {sorted(codes)}")
        self._chk(lib().mfc_set_bb_order(self.h, codes[order]))

    def set_timing(self, every):
        """Time every `every`-th collide launch with HIP events (True: each
        launch; False or 0: off)."""
        self._chk(lib().mfc_set_timing(self.h, int(every)))

    def get_timing(self):
        t = C.c_double()
        n = C.c_int64()
        self._chk(lib().mfc_get_timing(self.h, C.byref(t), C.byref(n)))
        return t.value, n.value


class Ensemble:
    """Many boxes advanced together (mfc_ens_*): one grid per step stage for
    all members.  Members are Lattice objects, ready to step (after macrovar,
    and beads_links when they hold spheres); they keep their own state."""

    def __init__(self):
        self.h = C.c_void_p()
        rc = lib().mfc_ens_create(C.byref(self.h))
        if rc != 0:
            raise MfcError(rc, "mfc_ens_create failed")
        self.members = []

    def _chk(self, rc):
        if rc != 0:
            raise MfcError(rc, lib().mfc_ens_last_error(self.h).decode())

    def add(self, lat: "Lattice", istep: int):
        self._chk(lib().mfc_ens_add(self.h, lat.h, int(istep)))
        self.members.append(lat)

    def remove(self, lat: "Lattice"):
        self._chk(lib().mfc_ens_remove(self.h, lat.h))
        self.members.remove(lat)

    def step(self, nsteps: int = 1):
        self._chk(lib().mfc_ens_step(self.h, int(nsteps)))

    def sync(self):
        self._chk(lib().mfc_ens_sync(self.h))

    def set_block(self, bs: int):
        self._chk(lib().mfc_ens_set_block(self.h, int(bs)))

    def set_graph(self, on: bool):
        """Replay each step as a recorded HIP graph (see mfc_ens_set_graph)."""
        self._chk(lib().mfc_ens_set_graph(self.h, int(bool(on))))

    def set_tile_order(self, order: int):
        """Collide tile order: 0 launch order, 1 contiguous ranges per XCD."""
        self._chk(lib().mfc_ens_set_tile_order(self.h, int(order)))

    def autotune(self, steps: int = 5):
        """Times every (collide block of 64, 128, 256 threads, tile order 0/1)
        pair on the current members (which advance 12 * steps steps); keeps
        the fastest.  Returns ((block, order), {(block, order): ms/step})."""
        chosen = np.zeros(2, dtype=np.int32)
        ms = np.zeros(6)
        self._chk(lib().mfc_ens_autotune(self.h, int(steps), _ip(chosen), _dp(ms)))
        keys = [(b, o) for o in (0, 1) for b in (64, 128, 256)]
        return (int(chosen[0]), int(chosen[1])), dict(zip(keys, ms))

    def host_time(self):
        """(host ms spent building and enqueueing steps, steps) since the
        last call."""
        ms = np.zeros(1)
        n = C.c_int64()
        self._chk(lib().mfc_ens_host_time(self.h, _dp(ms), C.byref(n)))
        return float(ms[0]), int(n.value)

    def close(self):
        if self.h:
            lib().mfc_ens_destroy(self.h)
            self.h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

