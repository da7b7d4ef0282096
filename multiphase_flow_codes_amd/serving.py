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

    def footprint_bytes(self, precision: str) -> int:
        """Device memory of the case (populations A/B, moments, maps, links)."""
        t = 8 if precision == "fp64" else 4
        c = self.cells
        b = 2 * 19 * t * c + 4 * 8 * c + 2 * 4 * c + c // 4
        r = self.rad
        seg = (math.ceil(2 * (r + 2) + 1) + 4)
        nseg = (seg + 3) // 4
        b += self.npart * nseg * 4 * seg * seg * 16
        b += self.npart * int(4 / 3 * math.pi * (r + 3) ** 3) * 9 * 21
        return int(b)


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
                 chunk=100, block=None, graph=True):
        assert window % chunk == 0 and cap % chunk == 0
        self.precision, self.budget, self.window, self.tol = precision, budget_bytes, window, tol
        self.cap, self.chunk, self.block, self.graph = cap, chunk, block, graph

    def run(self, cases: list, clock=time.perf_counter) -> dict:
        pending = sorted(cases, key=lambda c: c.arrival)
        active = {}                  # Lattice -> CaseRecord
        done = []
        ens = Ensemble()
        ens.set_graph(self.graph)
        if self.block:
            ens.set_block(self.block)
        used = 0
        cells_steps = 0
        t0 = clock()
        try:
            while pending or active:
                now = clock() - t0
                while pending and pending[0].arrival <= now:
                    fp = pending[0].footprint_bytes(self.precision)
                    if used + fp > self.budget and active:
                        break        # wait for memory to free up
                    c = pending.pop(0)
                    lat = make_lattice(c, self.precision)
                    ens.add(lat, 1)
                    active[lat] = CaseRecord(case=c, admitted=clock() - t0)
                    used += fp
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
        return summarize(done, wall, cells_steps)


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
    included); per-case mean drag averaged over the last `avg_window` steps,
    sampled every `sample_every` steps."""
    lats = [make_lattice(c, precision) for c in cases]
    ens = Ensemble()
    ens.set_graph(graph)
    try:
        for lat in lats:
            ens.add(lat, 1)
        block, tune = ens.autotune(tune_steps)
        done = 3 * tune_steps
        ens.sync()
        first = done
        acc = np.zeros(len(cases))
        nacc = 0
        wall = 0.0      # stepping time only: the drag sampling is not timed
        while done < steps:
            k = min(sample_every, steps - done)
            t0 = clock()
            ens.step(k)
            ens.sync()
            wall += clock() - t0
            done += k
            if steps - done < avg_window and k == sample_every:
                acc += np.array([mean_drag(l, c.npart) for l, c in zip(lats, cases)])
                nacc += 1
        cells = sum(c.cells for c in cases)
        return {"block": block, "tune_ms_per_step": tune, "steps": steps, "graph": bool(graph),
                "ms_per_step": 1e3 * wall / max(done - first, 1),
                "mlups": cells * (done - first) / wall / 1e6 if wall > 0 else 0.0,
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
