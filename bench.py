"""Throughput of the D3Q19 MRT step with moving spheres on one MI355X.

Headline workload (one "step" = one pass of the reference time-loop body,
BOX/main.f90:172-188: FORCING, collision_MRT, beads_collision, streaming,
beads_lubforce, beads_move, beads_redistribute, beads_links, beads_filling,
macrovar): channel1024, a 1024^3 channel with half-way bounce-back walls on
the y faces, decaying, nu = 0.002, 20,000 spheres of radius 6 (density ratio
2.5) at least 8 nodes from the walls, synthetic solenoidal initial field
(k0 = 8, u_rms = 0.03).  State is resident in HBM before the timed region.

Prints ONE JSON line: MLUPS of the FP32-shifted path with two population
buffers (`value`), the plain FP64 path with one buffer updated in place
beside it (two FP64 buffers do not fit), the HBM roofline fraction of the
dominant kernel (collide-stream, timed with HIP events on its own stream)
and of the whole step, FP32-vs-FP64 accuracy after 500 steps with an
FP64-vs-perturbed-FP64 control, the 256^3 box with 300 spheres with its own
accuracy legs, the 96-box ensemble, the serving loop, and the FP64 C oracle
timed on a bounded sample (a 256^3 section of the channel) on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"fp32": 152, "fp32x": 152, "fp64": 304}   # 19 populations read + 19 written
# HBM bytes per k_collide dispatch, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on these workloads (FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the
# factors a plain 1 GiB device copy gives); in-place storage: the mean of its two
# alternating layouts
TRAFFIC_SOURCE = "profiles/hbm_traffic_box256_channel1024.md"
MEASURED_TRAFFIC = {("box256", "fp32", "ab"): 2.64561e9, ("box256", "fp64", "ab"): 5.17563e9,
                    ("box256", "fp32x", "ab"): 2.63475e9,
                    ("channel1024", "fp32", "ab"): 1.63341e11, ("channel1024", "fp64", "aa"): 3.29672e11}
KERNEL = {"fp32": "k_collide<float,float>", "fp32x": "k_collide<float,double>",
          "fp64": "k_collide<double,double>"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(case, precision, steps, warmup, field, case_name, storage="ab", total_steps=0, collect=None,
            perturb=0.0):
    """Set-up, `warmup` untimed steps, `steps` timed steps (none when 0);
    then (accuracy legs) untimed steps up to `total_steps` and collect(lat)
    of the state."""
    from multiphase_flow_codes_amd.workloads import setup_lattice
    t0 = time.time()
    lat, prm, field = setup_lattice(case, precision, field, storage=storage, perturb=perturb)
    tag = (f"{case_name} {precision}{'' if storage == 'ab' else ' ' + storage.upper()}"
           f"{' perturbed' if perturb else ''}")
    log(f"{tag}: set-up {time.time() - t0:.1f} s")
    out = {"state": None}
    done = 0
    if steps > 0:
        lat.step(1, warmup)
        lat.sync()
        log(f"{tag}: warm-up {warmup} steps done")
        # HIP events around a sample of ~25 collide launches of the timed steps
        # (each event marker idles the GPU for a few microseconds); an odd period,
        # so in-place storage samples both of its alternating layouts
        every = max(1, steps // 25)
        lat.set_timing(every if every % 2 else every + 1)
        lat.sync()
        t1 = time.perf_counter()
        lat.step(1 + warmup, steps)
        lat.sync()
        t2 = time.perf_counter()
        kern_ms, launches = lat.get_timing()
        lat.set_timing(False)
        st = lat.status()
        ms_step = (t2 - t1) * 1e3 / steps
        ncell = prm.ncell
        mlups = ncell * steps / (t2 - t1) / 1e6
        avg_kern = kern_ms / max(launches, 1)
        bpc = BYTES_PER_CELL[precision]
        achieved = bpc * ncell / (avg_kern * 1e-3) / 1e9
        step_gbs = bpc * ncell / (ms_step * 1e-3) / 1e9
        log(f"{tag}: {ms_step:.3f} ms/step, {mlups:.0f} MLUPS, collide {avg_kern:.3f} ms "
            f"({achieved:.0f} GB/s), step {step_gbs:.0f} GB/s, links {st['nlinks']}, "
            f"contact links {st['contact_links']}, refilled nodes {st['refill']}")
        traffic = MEASURED_TRAFFIC.get((case_name, precision, storage))
        out.update({
            "mlups": mlups, "ms_per_step": ms_step, "kernel_ms": avg_kern, "launches": launches,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": traffic, "traffic_source": TRAFFIC_SOURCE if traffic else None,
                         "kernel": KERNEL[precision] + ("" if storage == "ab" else " (in-place AA layouts)"),
                         "bytes_per_cell": bpc,
                         "step_achieved": round(step_gbs, 1), "step_frac": round(step_gbs / HBM_PEAK_GBS, 4)},
            "links": st["nlinks"], "contact_links": st["contact_links"], "refilled": st["refill"]})
        done = warmup + steps
    if collect is not None:
        if total_steps > done:
            lat.step(1 + done, total_steps - done)
            lat.sync()
        out["state"] = collect(lat)
        out["contact_links_total"] = lat.status()["contact_links"]
        log(f"{tag}: state after {max(total_steps, done)} steps collected")
    lat.close()
    return out, field


def drag_report(Fa, Fb, pos=None, shape=None, rad=None, walls_y=False):
    """Per-sphere hydrodynamic force Fa against the reference Fb: normwise
    L2, the per-sphere error relative to the sphere's own |F| and to the rms
    |F|, and the worst sphere with its surface gap to its nearest neighbour
    (and to a wall)."""
    import numpy as np
    fn = np.linalg.norm(Fb, axis=1)
    dfn = np.linalg.norm(Fa - Fb, axis=1)
    frms = float(np.sqrt(np.mean(fn ** 2)))
    rel = dfn / np.maximum(fn, 1e-300)
    k = int(np.argmax(rel))
    out = {"drag_rel_l2": float(np.linalg.norm(Fa - Fb) / np.linalg.norm(Fb)),
           "drag_max_rel": float(rel.max()), "drag_median_rel": float(np.median(rel)),
           "drag_max_err_over_rms_drag": float(dfn.max() / frms),
           "worst_sphere": {"id": k + 1, "rel_err": float(rel[k]), "abs_F_over_rms": float(fn[k] / frms)}}
    if pos is not None:
        d = pos - pos[k]
        L = np.array(shape, dtype=float)[::-1]
        for a in range(3):
            if a == 1 and walls_y:
                continue
            d[:, a] -= L[a] * np.round(d[:, a] / L[a])
        dist = np.linalg.norm(d, axis=1)
        dist[k] = np.inf
        out["worst_sphere"]["min_surface_gap"] = float(dist.min() - 2 * rad)
        if walls_y:
            out["worst_sphere"]["wall_gap"] = float(min(pos[k, 1], L[1] - pos[k, 1]) - rad)
    return out


def run_accuracy(case, field, nsteps, modes=("fp32",), control=6e-8):
    """FP32-storage paths against the plain FP64 path (the drop-in, which
    matches the FP64 oracle to <= 1e-12 per step) after nsteps from the same
    state: velocity relative L2, the statistc energy spectrum and the
    per-particle hydrodynamic force.  With `control` > 0, an FP64 run from an
    initial field perturbed by that relative amount (the size of one FP32
    rounding) is compared the same way: how far the flow itself carries a
    perturbation of that size."""
    import numpy as np
    from multiphase_flow_codes_amd.workloads import setup_lattice
    out = {}
    runs = [("fp64", 0.0)] + [(m, 0.0) for m in modes] + ([("fp64", control)] if control > 0 else [])
    for prec, pert in runs:
        key = prec if pert == 0 else "fp64_perturbed"
        lat, prm, _ = setup_lattice(case, prec, field if pert == 0 else None, perturb=pert)
        lat.step(1, nsteps)
        lat.sync()
        _, ux, uy, uz = lat.download_macro()
        nek = case.n // 2 - 1
        out[key] = {"u": np.stack([ux, uy, uz]), "ek": lat.diag_spectrum(nek),
                    "f": lat.prt_get("fHIp")[: case.npart], "pos": lat.prt_get("ypglb")[: case.npart],
                    "sums": lat.diag_sums(), "contact": lat.status()["contact_links"]}
        lat.close()
        log(f"accuracy: {key} {nsteps} steps done")
    b = out["fp64"]
    res = {}
    for mode in [m for m in out if m != "fp64"]:
        a = out[mode]
        du = float(np.linalg.norm(a["u"] - b["u"]) / np.linalg.norm(b["u"]))
        ek_rel = np.abs(a["ek"] - b["ek"]) / np.maximum(b["ek"], 1e-300)
        res[mode] = {
            "steps": nsteps, "reference": "plain FP64 GPU path",
            "velocity_rel_l2": du,
            "spectrum_max_rel": float(ek_rel.max()),
            "spectrum_energy_weighted_rel": float(np.abs(a["ek"] - b["ek"]).sum() / b["ek"].sum()),
            "kinetic_energy_rel": float(abs(a["sums"]["u2"] - b["sums"]["u2"]) / b["sums"]["u2"]),
            "contact_links_total": a["contact"]}
        res[mode].update(drag_report(a["f"], b["f"], b["pos"], case.shape, case.rad, case.walls))
        if mode == "fp64_perturbed":
            res[mode]["perturbation"] = f"initial mode amplitudes x (1 + {control:g} xi), xi ~ N(0,1)"
        log(f"accuracy {mode}: {res[mode]}")
    return res


def run_ensemble_bench(ncases, steps):
    """Batched ensemble (mixed batch of small boxes, FP32) against the same
    ensemble in FP64: aggregate MLUPS, the autotuned collide pencil (block
    size and tile order), host time per step and per-case mean drag error."""
    import numpy as np
    from multiphase_flow_codes_amd import serving
    cases = serving.draw_cases(ncases, np.random.default_rng(2026))
    r32 = serving.run_ensemble(cases, steps, "fp32")
    log(f"ensemble fp32: {r32['mlups']:.0f} MLUPS, {r32['ms_per_step']:.3f} ms/step, block {r32['block']}, "
        f"tile order {r32['tile_order']}, host {r32['host_ms_per_step']:.3f} ms/step")
    r64 = serving.run_ensemble(cases, steps, "fp64")
    log(f"ensemble fp64: {r64['mlups']:.0f} MLUPS, {r64['ms_per_step']:.3f} ms/step, block {r64['block']}, "
        f"tile order {r64['tile_order']}, host {r64['host_ms_per_step']:.3f} ms/step")
    rel = np.abs(r32["mean_drag"] - r64["mean_drag"]) / np.maximum(np.abs(r64["mean_drag"]), 1e-300)
    keep = ("mlups", "ms_per_step", "host_ms_per_step", "block", "tile_order", "timed_steps")
    return {"workload": f"{ncases} periodic boxes, edges from {{32,48,64,96,128}}, 1-50 spheres r 3-8, "
                        f"solid fraction <= 10%, density ratio U[1.5,10], x body force log-U[1e-6,1e-4], "
                        f"nu=0.01, from rest, {steps} steps",
            "cells": int(sum(c.cells for c in cases)), "spheres": int(sum(c.npart for c in cases)),
            "step_launch": ("chunks of 8 steps: one table upload and one recorded HIP graph "
                            "(8 stage grids per step) per chunk") if r32.get("graph") else "8 stage grids per step",
            "pencil": "collide block = `block` consecutive x nodes of one box; tile order 1 = blocks dealt "
                      "to the 8 XCDs in contiguous ranges",
            "fp32": {k: (round(r32[k], 4) if isinstance(r32[k], float) else r32[k]) for k in keep}
                    | {"autotune_ms_per_step": {k: round(v, 4) for k, v in r32["tune_ms_per_step"].items()}},
            "fp64": {k: (round(r64[k], 4) if isinstance(r64[k], float) else r64[k]) for k in keep},
            "mean_drag_rel_err_vs_fp64": {"max": float(rel.max()), "median": float(np.median(rel))}}


def run_serving_bench(ncases, cap, rate, window=500, tol=1e-4):
    """Serving loop (continuous batching) in FP32 and in FP64 against
    one-at-a-time plain FP64 solves of the same cases under the same arrival
    times.  The FP64 engine's results are bitwise the one-at-a-time ones, so
    its cases/hour ratio is what batching alone gives; the FP32 engine adds
    the precision.  Drag differences at the same step are stated relative to
    each case's terminal drag (its spheres' buoyant weight)."""
    import numpy as np
    from multiphase_flow_codes_amd import serving
    rng = np.random.default_rng(31)
    cases = serving.draw_cases(ncases, rng, edges=tuple(range(32, 161, 16)), max_spheres=64)
    serving.poisson_arrivals(cases, rate, rng)
    runs = {}
    for prec in ("fp32", "fp64"):
        runs[prec] = serving.Server(prec, budget_bytes=240e9, window=window, tol=tol, cap=cap).run(cases)
        r = runs[prec]
        log(f"serving {prec}: p50 {r['p50_s']:.2f} s, p99 {r['p99_s']:.2f} s, {r['cases_per_hour']:.0f} cases/h, "
            f"peak footprint {r['peak_footprint_bytes'] / 1e9:.1f} GB, waits {r['admission_waits']}")
    # the same stream under a budget of a quarter of the FP32 engine's peak
    # footprint (at least the largest case): admission waits, scheduling
    # changes, results do not
    tight_budget = max(0.25 * runs["fp32"]["peak_footprint_bytes"],
                       max(c.footprint_bytes("fp32") for c in cases))
    tight = serving.Server("fp32", budget_bytes=tight_budget, window=window, tol=tol, cap=cap).run(cases)
    log(f"serving fp32 under {tight_budget / 1e9:.2f} GB: p50 {tight['p50_s']:.2f} s, p99 {tight['p99_s']:.2f} s, "
        f"{tight['cases_per_hour']:.0f} cases/h, peak footprint {tight['peak_footprint_bytes'] / 1e9:.2f} GB, "
        f"waits {tight['admission_waits']}")
    one = serving.run_one_at_a_time(cases, "fp64", window=window, tol=tol, cap=cap)
    log(f"one-at-a-time fp64: p50 {one['p50_s']:.2f} s, p99 {one['p99_s']:.2f} s, "
        f"{one['cases_per_hour']:.0f} cases/h")

    def same_step(a_hist, b_hist):
        rel, norm, differ, worst = [], [], 0, None
        for c in cases:
            a, b = a_hist[c.seed], b_hist[c.seed]
            n = min(len(a), len(b))
            differ += len(a) != len(b)
            if n:
                d = abs(a[n - 1] - b[n - 1])
                rel.append(d / max(abs(b[n - 1]), 1e-300))
                norm.append(d / (c.npart * c.terminal_drag))
                if worst is None or rel[-1] > worst["rel"]:
                    worst = {"rel": rel[-1], "over_terminal_drag": norm[-1],
                             "mean_drag_over_terminal": abs(b[n - 1]) / (c.npart * c.terminal_drag),
                             "edge": c.n, "spheres": c.npart, "step": window * n}
        return {"max_rel": float(np.max(rel)), "median_rel": float(np.median(rel)),
                "max_over_terminal_drag": float(np.max(norm)), "median_over_terminal_drag": float(np.median(norm)),
                "worst_case": worst, "cases_stopped_at_other_step": differ}

    keep = ("p50_s", "p99_s", "cases_per_hour", "mlups", "wall_s", "converged", "admission_waits",
            "peak_footprint_bytes", "max_active")
    fmt = lambda r: {k: round(r[k], 4) if isinstance(r[k], float) else r[k] for k in keep if k in r}
    return {"workload": f"{ncases} cases, Poisson arrivals {rate}/s, edges 32-160, <= 64 spheres, "
                        f"run until mean drag changes < {tol} over {window} steps, cap {cap} steps, "
                        f"240 GB admission budget",
            "engine_fp32": fmt(runs["fp32"]), "engine_fp64": fmt(runs["fp64"]),
            "engine_fp32_tight_budget": fmt(tight) | {
                "budget_bytes": tight_budget,
                "peak_within_budget": tight["peak_footprint_bytes"] <= tight_budget,
                "drag_bitwise_equal_to_unbounded_engine": all(
                    tight["drag_history"][c.seed] == runs["fp32"]["drag_history"][c.seed] for c in cases)},
            "one_at_a_time_fp64": fmt(one),
            "cases_per_hour_ratio_fp32_engine": round(runs["fp32"]["cases_per_hour"] / one["cases_per_hour"], 3),
            "cases_per_hour_ratio_fp64_engine": round(runs["fp64"]["cases_per_hour"] / one["cases_per_hour"], 3),
            "fp64_engine_drag_bitwise_equal": all(runs["fp64"]["drag_history"][c.seed] == one["drag_history"][c.seed]
                                                  for c in cases),
            "fp32_engine_drag_vs_fp64_same_step": same_step(runs["fp32"]["drag_history"], one["drag_history"])}


def _channel_state(lat):
    """Velocity (kept in FP32 on the host: 13 GB at 1024^3), the statistc
    spectrum and the per-sphere hydrodynamic force and position."""
    import numpy as np
    _, ux, uy, uz = lat.download_macro()
    u = [a.astype(np.float32) for a in (ux, uy, uz)]
    del ux, uy, uz
    return {"u": u, "ek": lat.diag_spectrum(lat.prm.nx // 2 - 1), "f": lat.prt_get("fHIp"),
            "pos": lat.prt_get("ypglb")}


def _sq_sum(arrs, other=None, chunk=1 << 26):
    """sum of squares (or of squared differences) in FP64, in chunks"""
    import numpy as np
    s = 0.0
    for k, a in enumerate(arrs):
        fa = a.reshape(-1)
        fb = None if other is None else other[k].reshape(-1)
        for i in range(0, fa.size, chunk):
            x = fa[i:i + chunk].astype(np.float64)
            if fb is not None:
                x -= fb[i:i + chunk]
            s += float(x @ x)
    return s


def run_channel(steps, warmup, fp32_aa=False, fp64=True, acc_steps=500, control=6e-8):
    """The 1024^3 channel with 20,000 spheres: FP32 shifted populations with
    two buffers (163 GB) beside FP64 with one buffer updated in place
    (AA pattern, 163 GB); two FP64 buffers would need 326 GB.  With acc_steps
    both runs continue to that step and the FP32 velocity, spectrum and
    sphere drag are compared with the FP64 run's; with `control` an FP64 run
    from an initial field perturbed by that relative amount is compared the
    same way."""
    import numpy as np
    from multiphase_flow_codes_amd import workloads
    case = workloads.CHANNEL1024
    keep = ("mlups", "ms_per_step", "kernel_ms", "roofline", "links", "contact_links", "refilled")
    acc = acc_steps > 0 and fp64
    col = _channel_state if acc else None
    r32, _ = run_gpu(case, "fp32", steps, warmup, None, "channel1024", total_steps=acc_steps, collect=col)
    out = {"workload": case.describe(), "cells": case.n ** 3, "particles": case.npart,
           "steps": steps, "warmup": warmup,
           "fp32": {k: r32[k] for k in keep} | {"storage": "two buffers (A/B), 163 GB"}}
    if fp32_aa:
        ra, _ = run_gpu(case, "fp32", steps, warmup, None, "channel1024", storage="aa")
        out["fp32_aa"] = {k: ra[k] for k in keep} | {"storage": "one buffer in place (AA), 82 GB"}
    if fp64:
        r64, _ = run_gpu(case, "fp64", steps, warmup, None, "channel1024", storage="aa",
                         total_steps=acc_steps, collect=col)
        out["fp64"] = {k: r64[k] for k in keep} | {"storage": "one buffer in place (AA), 163 GB"}
        out["fp32_over_fp64"] = round(r32["mlups"] / r64["mlups"], 3)
    if acc:
        b = r64["state"]
        legs = [("accuracy_fp32_vs_fp64", r32["state"])]
        if control > 0:
            rp, _ = run_gpu(case, "fp64", 0, 0, None, "channel1024", storage="aa", total_steps=acc_steps,
                            collect=col, perturb=control)
            legs.append(("control_fp64_perturbed_vs_fp64", rp["state"]))
        for name, a in legs:
            du = float(np.sqrt(_sq_sum(a["u"], b["u"]) / _sq_sum(b["u"])))
            ek = np.abs(a["ek"] - b["ek"]) / np.maximum(b["ek"], 1e-300)
            out[name] = {
                "steps": acc_steps, "reference": "FP64 run of the same case (in-place storage)",
                "velocity_rel_l2": du, "note": "velocities compared as FP32 copies (6e-8 relative)",
                "spectrum_max_rel": float(ek.max())}
            out[name].update(drag_report(a["f"], b["f"], b["pos"], case.shape, case.rad, True))
            if name.startswith("control"):
                out[name]["perturbation"] = f"initial mode amplitudes x (1 + {control:g} xi), xi ~ N(0,1)"
            log(f"channel {name}: {out[name]}")
    return out


def cpu_info():
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def run_cpu_baseline(nsteps):
    """The FP64 C oracle (reference single-rank step order, OpenMP over the
    fluid loops) on channel256, a 256^3 section of the headline channel with
    its solid fraction, for nsteps steps from the same kind of initial state."""
    import numpy as np
    from oracle import pyoracle
    from multiphase_flow_codes_amd import workloads
    from multiphase_flow_codes_amd.initial import random_spheres
    import subprocess
    if not os.path.exists(pyoracle.LIB_PATH):
        subprocess.check_call(["bash", os.path.join(ROOT, "oracle", "build.sh")])
    case = workloads.CHANNEL256
    prm = case.params()
    o = pyoracle.Oracle(prm, bb_mode=0)
    o.ux[:], o.uy[:], o.uz[:] = case.host_field()
    o.initpop()
    o.rhoupdat()
    o.macrovar()
    pos = random_spheres(case.npart, case.rad, case.shape, gap=prm.mingap, wall_clear=case.wall_clear,
                         walls_y=True, seed=prm.iseedp)
    o.ypglb[:] = pos
    o.yp[:] = pos
    o.wp[:] = 0.0
    o.omgp[:] = 0.0
    o.S.istep00 = 1
    o.beads_links()
    nthreads = pyoracle.lib().orc_nthreads()
    t0 = time.perf_counter()
    for it in range(1, nsteps + 1):
        o.S.istep = it
        o.call("timestep")
    dt = time.perf_counter() - t0
    mlups = prm.ncell * nsteps / dt / 1e6
    log(f"cpu oracle: {nsteps} steps in {dt:.1f} s -> {mlups:.2f} MLUPS on {nthreads} threads")
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": nthreads, "kind": "port",
            "sample": f"{nsteps} steps ({dt:.1f} s) of {case.describe()} -- a 256^3 section of the "
                      f"headline channel with its solid fraction; FP64 C oracle in the reference's "
                      f"single-rank step order, OpenMP over the fluid loops"} | cpu_info()


def run_contact_orders(steps=20):
    """Both bounce-back orders (mfc_set_bb_order) on 1100 spheres of radius
    2.5 at surface gaps from 0.4 in a 96 x 88 x 104 periodic box, FP64, fluid
    and spheres starting at rest: contact links in the first steps, until
    lubrication separates the pairs (profiling: run under a kernel trace and
    compare k_bb_contact per step; the cumulative contact-link count is
    printed after every step)."""
    import numpy as np
    from multiphase_flow_codes_amd.initial import random_spheres
    from multiphase_flow_codes_amd.lbm import Lattice
    from multiphase_flow_codes_amd.params import LBMParams
    prm = LBMParams(nx=96, ny=88, nz=104, visc=0.01, npart=1100, rad=2.5, rhopart=1.5, forcing=False)
    pos = random_spheres(prm.npart, prm.rad, (prm.nz, prm.ny, prm.nx), gap=0.4, seed=17)
    zero = np.zeros((prm.nz, prm.ny, prm.nx))
    out = {}
    for order in ("snapshot", "reference"):
        lat = Lattice(prm, "fp64")
        try:
            lat.set_bb_order(order)
            lat.upload_macro(None, zero, zero, zero)
            lat.initpop()
            lat.rhoupdat()
            lat.macrovar()
            lat.prt_init(prm.npart)
            lat.prt_set("ypglb", pos)
            lat.initpartvel()
            lat.initpartomg()
            lat.beads_links()
            counts = []
            for it in range(1, steps + 1):
                lat.step(it, 1)
                lat.sync()
                counts.append(lat.status()["contact_links"])
            log(f"{order}: cumulative contact links per step {counts}")
            out[order] = counts
        finally:
            lat.close()
    return out


def calibrate_copy(mib):
    """A plain device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch against mib*1024 KiB gives the scale
    factor of those counters on this device."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: 5 copies of {mib} MiB done")


def hbm_ceilings(gib=4, reps=5):
    """Device bandwidth ceilings with simple kernels over `gib` GiB buffers:
    runtime copy, out-of-place elementwise (read one buffer, write another),
    in-place elementwise (read and write the same buffer), write-only fill."""
    import torch
    n = gib * (1 << 30) // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    nbytes = gib * (1 << 30)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    for name, fn, moved in (("copy", lambda: b.copy_(a), 2),
                            ("elementwise_out_of_place", lambda: torch.add(a, 1.0, out=b), 2),
                            ("elementwise_in_place", lambda: a.mul_(1.0000001), 2),
                            ("fill", lambda: b.fill_(0.5), 1)):
        ms = timed(fn)
        out[name] = round(moved * nbytes / ms / 1e6, 1)
        log(f"ceiling {name}: {ms:.3f} ms for {gib} GiB -> {out[name]:.0f} GB/s")
    print(json.dumps({"hbm_ceilings_GBps": out, "buffer_GiB": gib}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps of every throughput leg")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--channel-accuracy-steps", type=int, default=500,
                    help="run the channel legs to this step and compare FP32 with FP64 (0: skip)")
    ap.add_argument("--control", type=float, default=6e-8,
                    help="relative perturbation of the FP64 control runs (0: no control runs)")
    ap.add_argument("--channel-fp32-aa", action="store_true",
                    help="also run the 1024^3 channel in FP32 with in-place storage")
    ap.add_argument("--channel-no-fp64", action="store_true")
    ap.add_argument("--channel-only", action="store_true",
                    help="run only the 1024^3 channel legs (profiling); prints their JSON on stderr")
    ap.add_argument("--no-box256", action="store_true")
    ap.add_argument("--box256-only", action="store_true",
                    help="run only the 256^3 box legs (profiling); prints their JSON on stderr")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the 256^3 FP32-vs-FP64 accuracy comparison (0: skip)")
    ap.add_argument("--no-fp32x", action="store_true",
                    help="skip the FP32-storage / FP64-collision variant of the 256^3 box")
    ap.add_argument("--ens-cases", type=int, default=96)
    ap.add_argument("--ens-steps", type=int, default=2000)
    ap.add_argument("--serve-cases", type=int, default=24)
    ap.add_argument("--serve-cap", type=int, default=20000,
                    help="step cap per served case")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="arrivals per second")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the ensemble legs (profiling); prints their JSON on stderr")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized buffer 5 times (HBM counter calibration)")
    ap.add_argument("--ceilings", action="store_true",
                    help="only measure the device's bandwidth ceilings with simple kernels")
    ap.add_argument("--contact-orders", action="store_true",
                    help="only step 1100 close spheres in both bounce-back orders (profiling)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.ceilings:
        hbm_ceilings()
        return
    if args.contact_orders:
        log(json.dumps(run_contact_orders()))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; sharding over several GPUs "
                                   "(z-slab halos) is designed but not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.ensemble_only:
        log(json.dumps(run_ensemble_bench(args.ens_cases, args.ens_steps)))
        return
    box = None
    if not args.channel_only and not args.no_box256:
        box = run_box256(args)
        if args.box256_only:
            log(json.dumps(box))
            return
    c3 = run_channel(args.steps, args.warmup, args.channel_fp32_aa, not args.channel_no_fp64,
                     args.channel_accuracy_steps, args.control)
    if args.channel_only:
        log(json.dumps(c3))
        return
    ens = None if args.no_ensemble else run_ensemble_bench(args.ens_cases, args.ens_steps)
    srv = None if args.no_serving else run_serving_bench(args.serve_cases, args.serve_cap, args.serve_rate)
    cpu = None if args.no_cpu else run_cpu_baseline(args.cpu_steps)
    f32 = c3["fp32"]
    out = {
        "metric": "MLUPS + %HBM roofline, FP32 vs FP64 path; p50/p99 case latency (mixed batch)",
        "value": round(f32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": c3["workload"], "cells": c3["cells"], "particles": c3["particles"],
                   "boundary_links": f32["links"], "contact_links": f32["contact_links"],
                   "populations": "FP32 shifted deviations, two buffers (A/B)",
                   "particle_arithmetic": "f64"},
        "roofline": f32["roofline"],
        "cpu_baseline": cpu,
    }
    if "fp64" in c3:
        f64 = c3["fp64"]
        out["fp64"] = {"value": round(f64["mlups"], 1), "unit": "MLUPS",
                       "ms_per_step": round(f64["ms_per_step"], 4), "roofline": f64["roofline"],
                       "storage": f64["storage"], "contact_links": f64["contact_links"]}
        out["fp32_over_fp64"] = c3["fp32_over_fp64"]
    for k in ("accuracy_fp32_vs_fp64", "control_fp64_perturbed_vs_fp64", "fp32_aa"):
        if k in c3:
            out[k] = c3[k]
    if box is not None:
        out["box256"] = box
    if ens is not None:
        out["ensemble"] = ens
    if srv is not None:
        out["serving"] = srv
        out["p50_case_latency_s"] = srv["engine_fp32"]["p50_s"]
        out["p99_case_latency_s"] = srv["engine_fp32"]["p99_s"]
    print(json.dumps(out), flush=True)


def run_box256(args):
    """box256: the 256^3 periodic box with 300 spheres, FP32, FP64 and
    FP32-storage/FP64-collision throughput, and their accuracy against FP64
    after args.accuracy_steps steps with the perturbed-FP64 control."""
    from multiphase_flow_codes_amd import workloads
    case = workloads.BOX256
    keep = ("mlups", "ms_per_step", "kernel_ms", "roofline", "links", "contact_links", "refilled")
    r32, field = run_gpu(case, "fp32", args.steps, args.warmup, None, "box256")
    r64, field = run_gpu(case, "fp64", args.steps, args.warmup, field, "box256")
    out = {"workload": case.describe(), "cells": case.n ** 3, "particles": case.npart,
           "fp32": {k: r32[k] for k in keep}, "fp64": {k: r64[k] for k in keep},
           "fp32_over_fp64": round(r32["mlups"] / r64["mlups"], 3)}
    if not args.no_fp32x:
        r32x, field = run_gpu(case, "fp32x", args.steps, args.warmup, field, "box256")
        out["fp32_storage_fp64_collision"] = {k: r32x[k] for k in keep}
        out["fp32_storage_fp64_collision"]["over_fp64"] = round(r32x["mlups"] / r64["mlups"], 3)
    if args.accuracy_steps > 0:
        acc = run_accuracy(case, field, args.accuracy_steps,
                           ("fp32",) if args.no_fp32x else ("fp32", "fp32x"), args.control)
        out["accuracy_fp32_vs_fp64"] = acc["fp32"]
        if "fp32x" in acc:
            out["fp32_storage_fp64_collision"]["accuracy_vs_fp64"] = acc["fp32x"]
        if "fp64_perturbed" in acc:
            out["control_fp64_perturbed_vs_fp64"] = acc["fp64_perturbed"]
    return out


if __name__ == "__main__":
    main()
