"""Batched ensembles and a serving loop over the ensemble engine.

Cases are small periodic boxes of spheres driven by a uniform x body force
(the mixed-batch workloads: edges 32-128 or 32-160, 1-64 spheres of radius
3-8, solid fraction <= 10 %, density ratio 1.5-10, body force log-uniform in
[1e-6, 1e-4], nu = 0.01).  A periodic box has no momentum sink, so the drive
is posed as sedimentation with back-flow: the acceleration g acts on the
spheres' excess mass, (rho_p/rho_f - 1) g per unit sphere volume, and the
fluid carries the uniform counter-force -phi (rho_p/rho_f - 1) g (phi: solid
fraction), so the total external force is zero and the spheres reach a
terminal velocity, at which their mean drag balances their buoyant weight.
g is capped so that the Stokes settling speed stays below 0.02 (low Mach
number).  Each case is an ordinary Lattice; the engine
advances all admitted cases together (Ensemble, one grid per step stage), so
a case's state is bitwise the same whether it runs alone or in the mix.

The serving loop admits arriving cases while their HBM footprint fits the
budget, steps the admitted ones together, checks each case's mean particle
drag every `window` steps and retires it when the drag changed by less than
`tol` (relative) over the last window, or at `cap` steps.  It reports the
time from arrival to result of every case.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .lbm import Ensemble, Lattice
from .params import LBMParams


def device_free_bytes() -> int:
    """Free device memory of the current GPU (hipMemGetInfo)."""
    import ctypes as C
    import os
    global _hip
    if _hip is None:
        root = os.environ.get("ROCM_PATH", "/opt/rocm")
        _hip = C.CDLL(os.path.join(root, "lib", "libamdhip64.so"))
    free, total = C.c_size_t(), C.c_size_t()
    if _hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return int(free.value)


_hip = None


@dataclass
class BoxCase:
    n: int                      # edge length (cubic box)
    npart: int
    rad: float
    rho_ratio: float
    body_force: float           # x component
    seed: int
    visc: float = 0.01
    arrival: float = 0.0        # seconds after the start of a serving run

    @property
    def phi(self) -> float:
        return self.npart * 4.0 / 3.0 * math.pi * self.rad ** 3 / self.n ** 3

    def params(self) -> LBMParams:
        g = self.body_force
        excess = (self.rho_ratio - 1.0) * g
        return LBMParams(nx=self.n, ny=self.n, nz=self.n, visc=self.visc, npart=self.npart,
                         rad=self.rad, rhopart=self.rho_ratio, forcing=False,
                         body_force=(-self.phi * excess, 0.0, 0.0),
                         particle_force=(excess, 0.0, 0.0), iseedp=self.seed, irelease=1)

    @property
    def cells(self) -> int:
        return self.n ** 3

    @property
    def terminal_drag(self) -> float:
        """Magnitude of the mean drag at terminal velocity: the buoyant weight
        per sphere, V_p (rho_p/rho_f - 1) g."""
        return 4.0 / 3.0 * math.pi * self.rad ** 3 * (self.rho_ratio - 1.0) * self.body_force

    def footprint_bytes(self, precision: str) -> int:
        """Device memory make_lattice allocates for the case: the sizes
        mfc_create and mfc_prt_init request (populations A/B, 19 values per
        node each, forcing tables, moments, particle arrays, segment lists, link
        pool, refill list, node maps), each rounded up to the allocator's
        granularity (4 KiB below 1 MiB, 2 MiB above), plus 16 MiB for the
        runtime's per-context state.  An upper bound of the hipMemGetInfo
        difference around make_lattice (checked by the serving test)."""
        t = 8 if precision == "fp64" else 4
        n, np_, r = self.n, self.npart, self.rad
        ncell = n ** 3
        allocs = [19 * ncell * t] * 2 + [8 * 18 * n * n] * 4 + [8 * 6 * n, 64, 64] + [8 * ncell] * 4
        if np_ > 0:
            allocs += [8 * 3 * np_] * 16
            maxcnt = math.ceil(2.0 * (r + 2.0) + 1.0) + 4
            nseg, seg_cap = (maxcnt + 3) // 4, 4 * maxcnt * maxcnt
            nsegs = np_ * nseg
            nodes = nsegs * seg_cap
            allocs += [4 * nsegs, 8 * (nsegs + 1), 16 * (nsegs // 4096 + 1)]
            allocs += [4 * nodes, 4 * nsegs] * 2 + [4 * nodes, 4 * nodes, 4 * nsegs, 8 * 6 * nsegs * 4]
            allocs += [8 * (math.ceil(4.0 / 3.0 * math.pi * (r + 1.0) ** 3) * np_ + 64), 4]
            h = math.sqrt(3.0) / 2.0
            ro, ri = r + math.sqrt(2.0) + h, max(0.0, r - h)
            cap = (math.ceil(4.0 / 3.0 * math.pi * (ro ** 3 - ri ** 3)) * 9 + 64) * np_
            allocs += [4, 4 * cap, cap, 8 * cap, 16 * cap, 8 * cap]
            allocs += [4 * ncell, 4 * ((ncell + 31) // 32)] * 2
        total = 16 << 20
        for x in allocs:
            g = (2 << 20) if x >= (1 << 20) else 4096
            total += (x + g - 1) // g * g
        return int(total)


def draw_cases(count: int, rng: np.random.Generator, edges=(32, 48, 64, 96, 128), max_spheres=50,
               rad=(3.0, 8.0), phi_max=0.10, rho=(1.5, 10.0), force=(1e-6, 1e-4)) -> list:
    """Cases of the mixed batch: the edge drawn from `edges`, radius and
    density ratio uniform, body force log-uniform, sphere count uniform in
    1..min(max_spheres, the count that fills phi_max)."""
    out = []
    for k in range(count):
        n = int(rng.choice(edges))
        r = float(rng.uniform(*rad))
        r = min(r, float((n - 6) // 2 - 2))           # the box must hold one sphere's scan box
        nmax = int(phi_max * n ** 3 / (4.0 / 3.0 * math.pi * r ** 3))
        npart = int(rng.integers(1, max(1, min(max_spheres, nmax)) + 1))
        rho_ratio = float(rng.uniform(*rho))
        g = float(math.exp(rng.uniform(math.log(force[0]), math.log(force[1]))))
        # Stokes settling speed (rho_p - 1) g V_p / (6 pi nu r) kept below 0.02
        vol = 4.0 / 3.0 * math.pi * r ** 3
        g = min(g, 0.02 * 6.0 * math.pi * 0.01 * r / ((rho_ratio - 1.0) * vol))
        out.append(BoxCase(n=n, npart=npart, rad=r, rho_ratio=rho_ratio, body_force=g,
                           seed=int(rng.integers(1, 2 ** 30))))
    return out


def make_lattice(case: BoxCase, precision: str) -> Lattice:
    """The case at rest (fluid and spheres), spheres placed by the reference's
    seeded generator with the minimum gap, links built: ready to step from 1."""
    prm = case.params()
    lat = Lattice(prm, precision)
    z = np.zeros((prm.nz, prm.ny, prm.nx))
    lat.upload_macro(None, z, z, z)
    lat.initpop()
    lat.rhoupdat()
    lat.macrovar()
    if case.npart:
        lat.prt_init(case.npart)
        lat.initpartpos(case.seed, mingap_check=True)
        lat.initpartomg()
        lat.prt_set("wp", np.zeros((case.npart, 3)))
        lat.beads_links()
    lat.sync()
    return lat


def mean_drag(lat: Lattice, npart: int) -> float:
    """Mean x force of the fluid on the spheres (fHIp)."""
    if npart == 0:
        return 0.0
    return float(lat.prt_get("fHIp")[:npart, 0].mean())


@dataclass
class CaseRecord:
    case: BoxCase
    admitted: float = 0.0
    done: float = 0.0
    steps: int = 0
    drag: float = 0.0
    converged: bool = False
    drags: list = field(default_factory=list)

    @property
    def latency(self) -> float:
        return self.done - self.case.arrival


def _converged(rec: CaseRecord, tol: float) -> bool:
    if len(rec.drags) < 2:
        return False
    a, b = rec.drags[-2], rec.drags[-1]
    return abs(b - a) <= tol * max(abs(b), 1e-300)


class Server:
    """Continuous batching over one Ensemble."""

    def __init__(self, precision="fp32", budget_bytes=240e9, window=500, tol=1e-4, cap=20000,
                 chunk=100, block=None, graph=True, measure_memory=False):
        assert window % chunk == 0 and cap % chunk == 0
        self.precision, self.budget, self.window, self.tol = precision, budget_bytes, window, tol
        self.cap, self.chunk, self.block, self.graph = cap, chunk, block, graph
        self.measure_memory = measure_memory

    def run(self, cases: list, clock=time.perf_counter) -> dict:
        pending = sorted(cases, key=lambda c: c.arrival)
        active = {}                  # Lattice -> CaseRecord
        done = []
        ens = Ensemble()
        ens.set_graph(self.graph)
        if self.block:
            ens.set_block(self.block)
        used = 0
        peak = 0
        waits = 0
        max_active = 0
        mem = []                     # (estimated, measured) bytes of admitted cases
        cells_steps = 0
        t0 = clock()
        try:
            while pending or active:
                now = clock() - t0
                while pending and pending[0].arrival <= now:
                    fp = pending[0].footprint_bytes(self.precision)
                    if used + fp > self.budget and active:
                        waits += 1
                        break        # wait for memory to free up
                    c = pending.pop(0)
                    # measured while nothing has been freed yet (no reuse of freed memory)
                    measure = self.measure_memory and not done
                    free0 = device_free_bytes() if measure else 0
                    lat = make_lattice(c, self.precision)
                    if measure:
                        mem.append((fp, free0 - device_free_bytes()))
                    ens.add(lat, 1)
                    active[lat] = CaseRecord(case=c, admitted=clock() - t0)
                    used += fp
                    peak = max(peak, used)
                    max_active = max(max_active, len(active))
                if not active:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                    continue
                ens.step(self.chunk)
                cells_steps += self.chunk * sum(r.case.cells for r in active.values())
                retire = []
                for lat, rec in active.items():
                    rec.steps += self.chunk
                    if rec.steps % self.window == 0:
                        rec.drags.append(mean_drag(lat, rec.case.npart))
                        rec.converged = _converged(rec, self.tol)
                    if rec.converged or rec.steps >= self.cap:
                        retire.append(lat)
                for lat in retire:
                    rec = active.pop(lat)
                    rec.drag = rec.drags[-1] if rec.drags else mean_drag(lat, rec.case.npart)
                    ens.remove(lat)
                    rec.done = clock() - t0
                    lat.close()
                    used -= rec.case.footprint_bytes(self.precision)
                    done.append(rec)
            ens.sync()
        finally:
            for lat in list(active):
                ens.remove(lat)
                lat.close()
            ens.close()
        wall = clock() - t0
        out = summarize(done, wall, cells_steps)
        out.update({"admission_waits": waits, "peak_footprint_bytes": peak, "max_active": max_active,
                    "budget_bytes": self.budget})
        if self.measure_memory:
            out["footprint_estimated_measured"] = mem
        return out


def run_one_at_a_time(cases: list, precision="fp64", window=500, tol=1e-4, cap=20000, chunk=100,
                      clock=time.perf_counter) -> dict:
    """The same cases solved one after another with the single-box path: each
    case's service time is measured; arrival-to-result times follow from a
    first-come first-served queue over the same arrival times."""
    recs = []
    cells_steps = 0
    t_busy = 0.0
    for c in sorted(cases, key=lambda c: c.arrival):
        t0 = clock()
        lat = make_lattice(c, precision)
        rec = CaseRecord(case=c)
        it = 1
        while True:
            lat.step(it, chunk)
            it += chunk
            rec.steps += chunk
            if rec.steps % window == 0:
                rec.drags.append(mean_drag(lat, c.npart))
                rec.converged = _converged(rec, tol)
            if rec.converged or rec.steps >= cap:
                break
        lat.sync()
        rec.drag = rec.drags[-1] if rec.drags else mean_drag(lat, c.npart)
        lat.close()
        service = clock() - t0
        start = max(c.arrival, t_busy)
        rec.admitted = start
        rec.done = start + service
        t_busy = rec.done
        cells_steps += rec.steps * c.cells
        recs.append(rec)
    wall = max(r.done for r in recs) if recs else 0.0
    return summarize(recs, wall, cells_steps)


def summarize(recs: list, wall: float, cells_steps: int) -> dict:
    lat = np.array([r.latency for r in recs]) if recs else np.zeros(1)
    return {
        "cases": len(recs),
        "wall_s": wall,
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "cases_per_hour": 3600.0 * len(recs) / wall if wall > 0 else 0.0,
        "mlups": cells_steps / wall / 1e6 if wall > 0 else 0.0,
        "converged": int(sum(r.converged for r in recs)),
        "steps": [r.steps for r in recs],
        "drag": {r.case.seed: r.drag for r in recs},
        "drag_history": {r.case.seed: list(r.drags) for r in recs},   # every `window` steps
    }


def run_ensemble(cases: list, steps: int, precision="fp32", tune_steps=5, sample_every=10,
                 avg_window=500, graph=True, clock=time.perf_counter) -> dict:
    """All cases in one ensemble for `steps` steps (the autotuner's steps
    included).  The steps between the autotuner and the last `avg_window`
    are timed as one block (one synchronisation at its end); over the last
    `avg_window` steps the per-case mean drag is averaged, sampled every
    `sample_every` steps (untimed)."""
    lats = [make_lattice(c, precision) for c in cases]
    ens = Ensemble()
    ens.set_graph(graph)
    try:
        for lat in lats:
            ens.add(lat, 1)
        shape, tune = ens.autotune(tune_steps)
        done = 12 * tune_steps
        ens.sync()
        timed = max(0, steps - done - avg_window)
        ens.host_time()
        t0 = clock()
        ens.step(timed)
        host_ms, host_steps = ens.host_time()
        ens.sync()
        wall = clock() - t0
        done += timed
        acc = np.zeros(len(cases))
        nacc = 0
        while done < steps:
            k = min(sample_every, steps - done)
            ens.step(k)
            done += k
            if k == sample_every:
                acc += np.array([mean_drag(l, c.npart) for l, c in zip(lats, cases)])
                nacc += 1
        ens.sync()
        cells = sum(c.cells for c in cases)
        return {"block": shape[0], "tile_order": shape[1],
                "tune_ms_per_step": {f"{b}/{o}": v for (b, o), v in tune.items()},
                "steps": steps, "timed_steps": timed, "graph": bool(graph),
                "ms_per_step": 1e3 * wall / max(timed, 1),
                "host_ms_per_step": host_ms / max(host_steps, 1),
                "mlups": cells * timed / wall / 1e6 if wall > 0 else 0.0,
                "mean_drag": acc / max(nacc, 1)}
    finally:
        for lat in lats:
            if lat in ens.members:
                ens.remove(lat)
        ens.close()
        for lat in lats:
            lat.close()


def poisson_arrivals(cases: list, rate_per_s: float, rng: np.random.Generator) -> list:
    t = 0.0
    for c in cases:
        t += float(rng.exponential(1.0 / rate_per_s))
        c.arrival = t
    return cases
