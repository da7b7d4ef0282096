"""Throughput of the D3Q19 MRT step with moving spheres on one MI355X.

Workload (one "step" = one pass of the reference time loop body,
BOX/main.f90:172-188: FORCING, collision_MRT, beads_collision, streaming,
beads_lubforce, beads_move, beads_redistribute, beads_links, beads_filling,
macrovar): config 2, a 256^3 periodic box, nu = 0.0036, forcing on the
|k| <= sqrt(2) shell, 300 spheres of radius 8 (density ratio 5), synthetic
solenoidal initial field (k0 = 4, u_rms = 0.03).  State is resident in HBM
before the timed region starts.

Prints ONE JSON line: MLUPS of the FP32-shifted path (`value`), the plain
FP64 path beside it, the HBM roofline fraction of the dominant kernel
(collide-stream, timed with HIP events on its own stream) and the FP64 C
oracle timed on a bounded sample of the same case on the host cores.
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
BYTES_PER_CELL = {"fp32": 152, "fp64": 304}   # 19 populations read + 19 written
# HBM bytes per k_collide dispatch, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this workload (FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the
# factors a plain 1 GiB device copy gives): profiles/hbm_traffic_config2_256cube_300spheres.md
MEASURED_TRAFFIC = {("config2", "fp32"): 2.60276e9, ("config2", "fp64"): 5.15306e9}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(case, precision, steps, warmup, field, case_name):
    from multiphase_flow_codes_amd.workloads import setup_lattice
    t0 = time.time()
    lat, prm, field = setup_lattice(case, precision, field)
    log(f"{precision}: set-up {time.time() - t0:.1f} s")
    lat.step(1, warmup)
    lat.sync()
    log(f"{precision}: warm-up {warmup} steps done")
    lat.set_timing(True)
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
    achieved = BYTES_PER_CELL[precision] * ncell / (avg_kern * 1e-3) / 1e9
    log(f"{precision}: {ms_step:.3f} ms/step, {mlups:.0f} MLUPS, collide {avg_kern:.3f} ms "
        f"({achieved:.0f} GB/s), links {st['nlinks']}, refilled nodes {st['refill']}")
    lat.close()
    return {
        "mlups": mlups, "ms_per_step": ms_step, "kernel_ms": avg_kern, "launches": launches,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((case_name, precision)),
                     "kernel": f"k_collide<{ 'float' if precision == 'fp32' else 'double'}>",
                     "bytes_per_cell": BYTES_PER_CELL[precision]},
        "links": st["nlinks"],
    }, field


def run_cpu_baseline(case, field, nsteps):
    from multiphase_flow_codes_amd.workloads import BoxCase  # noqa: F401
    from oracle import pyoracle
    import subprocess
    if not os.path.exists(pyoracle.LIB_PATH):
        subprocess.check_call(["bash", os.path.join(ROOT, "oracle", "build.sh")])
    prm = case.params()
    o = pyoracle.Oracle(prm, bb_mode=0)
    o.ux[:], o.uy[:], o.uz[:] = field
    o.initpop()
    o.rhoupdat()
    o.macrovar()
    o.S.istep00 = 0
    o.initpartpos()
    o.S.istep00 = 1
    o.wp[:] = 0.0
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
            "sample": f"{nsteps} steps of the same {case.n}^3 / {case.npart}-sphere case, FP64 C "
                      f"oracle (reference single-rank order), OpenMP on the fluid loops"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--case", default="config2", choices=["config1", "config2"])
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-fp32", action="store_true",
                    help="skip the FP32 path (profiling runs of the FP64 kernels only)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized buffer 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a "
                                   "design note (z-slab halos), not an implemented path",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from multiphase_flow_codes_amd import workloads
    case = workloads.CONFIG2 if args.case == "config2" else workloads.CONFIG1
    r32 = r64 = field = None
    if not args.no_fp32:
        r32, field = run_gpu(case, "fp32", args.steps, args.warmup, None, args.case)
    if not args.no_fp64:
        r64, field = run_gpu(case, "fp64", args.steps, args.warmup, field, args.case)
    if r32 is None:
        log("FP32 path skipped: no headline line printed")
        return
    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(case, field, args.cpu_steps)
    out = {
        "metric": "MLUPS + %HBM roofline, FP32 vs FP64 path; p50/p99 case latency (mixed batch)",
        "value": round(r32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": case.describe(), "cells": case.n ** 3, "particles": case.npart,
                   "boundary_links": r32["links"], "populations": "FP32 shifted deviations",
                   "particle_arithmetic": "f64"},
        "roofline": r32["roofline"],
        "cpu_baseline": cpu,
    }
    if r64 is not None:
        out["fp64"] = {"value": round(r64["mlups"], 1), "unit": "MLUPS",
                       "ms_per_step": round(r64["ms_per_step"], 4), "roofline": r64["roofline"]}
        out["fp32_over_fp64"] = round(r32["mlups"] / r64["mlups"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
