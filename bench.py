#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the plain single-run GPU path, FP64, 1 GPU.

Workload (default): the Ceres H2O ballistic exosphere, physical mode - icosphere
642 vertices / 1280 faces, T = 100 + 135 max(cos sza, 0)^0.25, launch weighted
by Antoine sublimation flux, flux-Maxwellian emission, FP64 Stormer-Verlet with
dt = 10 s and at most 20000 steps, photolysis lifetime 6.4e5 s, freeze-out below
110 K, grid 50 exponential shells R..20R x 36 x 72 with density and per-cell
velocity moments - at 1e6 particles (synthetic inputs, seeded).

A "step" of this script is one complete run of that workload with its inputs
already resident in device memory (dsmc_run_create stages them once).  The line
printed is the JSON contract of the benchmark driver, with the roofline of the
particle kernel (HIP-event timing on the kernel's own stream) and the timed CPU
baseline (the oracle, OpenMP on the host cores, on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from dsmc_suite_amd import capi  # noqa: E402
from dsmc_suite_amd.capi import PRESET_CERES_PHYSICAL, Buffers, SurfaceArrays, ptr  # noqa: E402

METRIC = "particle-steps/sec FP64, plain vs batched/served, 1 MI355X; p50/p99 run latency"
# Algorithmic FP64 operations of one particle-step of the push (Verlet 51 +
# cell/photolysis/deposit bookkeeping 15), counted as in the design notes.
FLOPS_PER_STEP = 66
FP64_PEAK_TFLOPS = 78.6


def icosphere(lib, levels, lon=0.0, mass=2.99e-26, kB=1.380649e-23):
    nv, na, nf = C.c_int32(), C.c_int32(), C.c_int32()
    lib.dsmc_icosphere_sizes(levels, C.byref(nv), C.byref(na), C.byref(nf))
    vert = np.zeros(nv.value * 3)
    faces = np.zeros(na.value * 3, np.int32)
    fT, fw = np.zeros(nf.value), np.zeros(nf.value)
    rc = lib.dsmc_icosphere_build(levels, lon, 100.0, 135.0, mass, kB, ptr(vert, C.c_double),
                                  ptr(faces, C.c_int32), ptr(fT, C.c_double), ptr(fw, C.c_double))
    if rc:
        raise RuntimeError(f"dsmc_icosphere_build: {rc}")
    return SurfaceArrays(levels=levels, vert=vert, faces=faces, face_T=fT, face_w=fw)


def cpu_baseline(p, surf, sample: int, threads: int) -> dict:
    """The oracle (libm build, OpenMP) on the first `sample` particles."""
    from tests.helpers import oracle
    lib = oracle(devmath=False)
    q = p.copy()
    q.nparticles = sample
    b = Buffers(q, 5, per_particle=False, moments=True)
    t0 = time.perf_counter()
    rc = lib.cpu_ref_ballistic(C.byref(q), C.byref(surf.surface), C.byref(b.result), threads, 0)
    dt = time.perf_counter() - t0
    if rc:
        raise RuntimeError(f"cpu_ref_ballistic: {rc}")
    steps = b.result.total_steps
    return {"value": steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} particles of the same workload ({steps} particle-steps, "
                      f"{dt:.2f} s, oracle/libcpu_ref.so, OpenMP schedule(dynamic,100))"}


def plain_run(lib, surf, n, args, shape) -> dict:
    """Stages one plain run of n particles and times args.steps complete executions."""
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=n, seed=args.seed)
    outs = Buffers(p, 5, per_particle=False, moments=not args.no_moments)
    st = C.c_int(0)
    run = lib.dsmc_run_create(C.byref(p), C.byref(surf.surface), C.byref(outs.result), C.byref(st))
    if not run:
        raise RuntimeError(f"dsmc_run_create failed: {st.value}")
    shape_p = C.byref(shape) if shape is not None else None
    for _ in range(args.warmup):
        if lib.dsmc_run_launch(run, shape_p) or lib.dsmc_run_sync(run):
            raise RuntimeError("warmup launch failed")
    kms, fms = [], []
    rounds, fl = C.c_int(0), C.c_double(0)
    lib.dsmc_run_sync(run)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if lib.dsmc_run_launch(run, shape_p) or lib.dsmc_run_sync(run):
            raise RuntimeError("timed launch failed")
        kms.append(lib.dsmc_run_kernel_ms(run))
        lib.dsmc_run_timing(run, C.byref(fl), C.byref(rounds))
        fms.append(fl.value)
    lib.dsmc_run_sync(run)
    wall = time.perf_counter() - t0
    if lib.dsmc_run_collect(run, C.byref(outs.result)):
        raise RuntimeError("collect failed")
    lib.dsmc_run_destroy(run)
    steps_per_run = outs.result.total_steps
    favg = float(np.mean(fms)) / 1e3          # flight-kernel device time per run
    achieved = FLOPS_PER_STEP * steps_per_run / favg / 1e12
    return {"params": p, "value": steps_per_run * args.steps / wall, "ms_per_run": wall / args.steps * 1e3,
            "particles": n, "particle_steps_per_run": steps_per_run,
            "fates": {k: int(getattr(outs.result, k)) for k in
                      ("n_escaped", "n_frozen", "n_photolysed", "n_capped")},
            "run_device_ms": [round(x, 3) for x in kms],
            "flight_kernel_ms": [round(x, 3) for x in fms], "rounds": rounds.value,
            "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                         "flops_per_particle_step": FLOPS_PER_STEP,
                         "kernel": "flight_kernel<1,0,false>", "kernel_avg_ms_per_run": favg * 1e3}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=10_000_000)
    ap.add_argument("--also-particles", type=int, default=1_000_000, help="second plain size (0: none)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--relaunch-batch", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=60000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-moments", action="store_true", help="density only (no velocity moments)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)

    lib = capi.load()
    surf = icosphere(lib, 3)
    shape = None
    if args.block or args.blocks_per_cu or args.relaunch_batch:
        shape = capi.LaunchShape(args.block or 256, args.blocks_per_cu or 4, args.relaunch_batch or 16)
    main_run = plain_run(lib, surf, args.particles, args, shape)
    p = main_run.pop("params")
    line = {
        "metric": METRIC, "value": main_run["value"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main_run["ms_per_run"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded icosphere surface and Philox draws)",
        "config": {"workload": f"plain single run, Ceres H2O ballistic exosphere (physical mode) at "
                               f"{args.particles:.0e} particles: icosphere 642/1280, dt 10 s, <=20000 steps, "
                               f"grid 50x36x72 exponential to 20R, density + velocity moments per cell",
                   **{k: v for k, v in main_run.items() if k not in ("value", "ms_per_run", "roofline")}},
        "roofline": main_run["roofline"],
    }
    if args.also_particles:
        other = plain_run(lib, surf, args.also_particles, args, shape)
        other.pop("params")
        line["config"][f"plain_{args.also_particles:.0e}"] = {
            "value": other["value"], "ms_per_run": other["ms_per_run"],
            "flight_kernel_ms": other["flight_kernel_ms"], "rounds": other["rounds"]}
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(p, surf, args.cpu_sample, threads)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
