#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the plain single-run GPU path, FP64, 1 GPU.

Workload (default): the Ceres H2O ballistic exosphere, physical mode - icosphere
642 vertices / 1280 faces, T = 100 + 135 max(cos sza, 0)^0.25, launch weighted
by Antoine sublimation flux, flux-Maxwellian emission, FP64 Stormer-Verlet with
dt = 10 s and at most 20000 steps, photolysis lifetime 6.4e5 s, freeze-out below
110 K, grid 50 exponential shells R..20R x 36 x 72 with density and per-cell
velocity moments - at 1e7 particles (synthetic inputs, seeded); also at 1e6.

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
import math
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from dsmc_suite_amd import capi  # noqa: E402
from dsmc_suite_amd import workloads as W  # noqa: E402
from dsmc_suite_amd.capi import PRESET_CERES_PHYSICAL, PRESET_COMPAT_BALLISTIC, Buffers, SurfaceArrays, ptr  # noqa: E402

METRIC = "particle-steps/sec FP64, plain vs batched/served, 1 MI355X; p50/p99 run latency"
# Algorithmic FP64 operations of one particle-step: the Verlet push 51, the
# cell/photolysis/deposit bookkeeping 15.
FLOPS_PER_STEP = 66
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the event-batched flight kernel on the 1e7-particle
# Ceres run with the default launch shape, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (per dispatch, averaged over its launches: two per run), each
# scaled by what the counter reads for a plain 1 GiB device-to-device copy
# (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00).  Only that configuration reports it;
# others give null.
PMC_TRAFFIC_1E7 = 3_047_254_624 + 14_516_544_624   # read + written, bytes per launch


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


def host_cpus() -> dict:
    """The host cores this process may use: its affinity set, the cgroup's CPU
    quota (cpu.max, in whole cores), physical cores within the affinity set
    (lscpu), and the CPU model."""
    aff = sorted(os.sched_getaffinity(0))
    info = {"affinity_cpus": len(aff), "quota_cpus": None, "physical_cores": None, "model": None}
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["quota_cpus"] = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = {tuple(l.split(",")[1:3]) for l in out.splitlines()
                 if l and not l.startswith("#") and int(l.split(",")[0]) in set(aff)}
        info["physical_cores"] = len(cores) or None
        m = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        info["model"] = next((l.split(":", 1)[1].strip() for l in m.splitlines() if l.startswith("Model name")),
                             None)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if info["model"] is None:
        try:
            info["model"] = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                                 if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
    return info


def cpu_threads(info: dict) -> int:
    """All the cores the process can run on at once: the physical cores of the
    affinity set, bounded by the cgroup quota."""
    n = info["physical_cores"] or info["affinity_cpus"]
    if info["quota_cpus"]:
        n = min(n, info["quota_cpus"])
    return max(1, n)


def cpu_baseline(p, surf, sample: int, threads: int, single_sample: int) -> dict:
    """The oracle (libm build, OpenMP schedule(dynamic,100)) on the first `sample`
    particles with `threads` threads, and on the first `single_sample` with one."""
    from tests.helpers import oracle
    lib = oracle(devmath=False)

    def timed(n: int, t: int):
        q = p.copy()
        q.nparticles = n
        b = Buffers(q, 5, per_particle=False, moments=True)
        t0 = time.perf_counter()
        rc = lib.cpu_ref_ballistic(C.byref(q), C.byref(surf.surface), C.byref(b.result), t, 0)
        dt = time.perf_counter() - t0
        if rc:
            raise RuntimeError(f"cpu_ref_ballistic: {rc}")
        return b.result.total_steps, dt

    steps, dt = timed(sample, threads)
    s1, d1 = timed(single_sample, 1)
    return {"value": steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} particles of the same workload ({steps} particle-steps, "
                      f"{dt:.2f} s, oracle/libcpu_ref.so, OpenMP schedule(dynamic,100))",
            "one_thread": {"value": s1 / d1, "sample": f"first {single_sample} particles ({s1} particle-steps, "
                                                       f"{d1:.2f} s)"}}


THERMAL_GRIDS = ((36, 72), (360, 720))


def thermal_depth_points(tp, ffd: int) -> int:
    """Depth grid of the thermal model (thermal_model.c:109-117): dt = period /
    ffd / substeps, dr = sqrt(0.2 dt k / (rho Cp)), ceil(depth / dr) points (at
    least 3): 467 at the reference's 72 longitudes, 1477 at 720."""
    dt = 9.07417 * 3600.0 / ffd / tp.n_substep
    dr = math.sqrt(0.1 * 2.0 * dt * tp.conductivity / tp.density / tp.Cp)
    return max(3, math.ceil(tp.depth / dr))


def thermal_run(lib) -> dict:
    """The thermal model (thermal_model.c thermal_solve, default parameters, 5
    substeps) on the device at the reference's 36x72 grid and at 360x720; wall
    time of dsmc_thermal_solve (scratch allocation, upload, kernel, result back).
    Algorithmic HBM bytes per cell: the initial profile written (8 h), then per
    substep the forward sweep reads the old profile and writes the eliminated
    right-hand side, the back substitution reads it and writes the new profile
    (32 h); the rate below is that traffic over the whole call's wall time, so a
    lower bound on the kernel's own bandwidth."""
    tp = capi.ThermalParams()
    lib.thermal_params_default(C.byref(tp))
    out = {"substeps": int(tp.n_substep)}
    for ftd, ffd in THERMAL_GRIDS:
        T = np.zeros(ftd * ffd)
        ms = []
        for _ in range(3):
            t0 = time.perf_counter()
            if lib.dsmc_thermal_solve(ftd, ffd, C.byref(tp), None, T.ctypes.data_as(C.POINTER(C.c_double))):
                raise RuntimeError("dsmc_thermal_solve failed")
            ms.append((time.perf_counter() - t0) * 1e3)
        h = thermal_depth_points(tp, ffd)
        traffic = (8 * h + tp.n_substep * 32 * h) * ftd * ffd
        out[f"{ftd}x{ffd}_ms"] = round(min(ms), 3)
        out[f"{ftd}x{ffd}_depth_points"] = h
        out[f"{ftd}x{ffd}_cells_per_s"] = ftd * ffd / (min(ms) / 1e3)
        out[f"{ftd}x{ffd}_hbm_bytes"] = traffic
        out[f"{ftd}x{ffd}_GB_per_s"] = round(traffic / (min(ms) / 1e3) / 1e9, 1)
    return out


def thermal_cpu(threads: int) -> dict:
    """The oracle's thermal model on the host cores, same grids."""
    from tests.helpers import oracle
    lib = oracle(devmath=False)
    tp = capi.ThermalParams()
    lib.cpu_ref_thermal_params_default(C.byref(tp))
    out = {"cores": threads}
    for ftd, ffd in THERMAL_GRIDS:
        T = np.zeros(ftd * ffd)
        t0 = time.perf_counter()
        lib.cpu_ref_thermal(ftd, ffd, C.byref(tp), None, T.ctypes.data_as(C.POINTER(C.c_double)), threads)
        out[f"{ftd}x{ffd}_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    return out


def bumpy_mesh(lib, levels: int, noise: float = 0.15, seed: int = 2):
    """An icosphere (unit radius) with radial noise: a body that shadows itself."""
    nv, na, nf = C.c_int32(), C.c_int32(), C.c_int32()
    lib.dsmc_icosphere_sizes(levels, C.byref(nv), C.byref(na), C.byref(nf))
    vert = np.zeros(nv.value * 3)
    faces = np.zeros(na.value * 3, np.int32)
    lib.dsmc_icosphere_build(levels, 0.0, 100.0, 135.0, 3e-26, 1.38e-23, ptr(vert, C.c_double),
                             ptr(faces, C.c_int32), None, None)
    vert = vert.reshape(-1, 3) * (1.0 + noise * np.random.default_rng(seed).uniform(-1, 1, (nv.value, 1)))
    return np.ascontiguousarray(vert), np.ascontiguousarray(faces.reshape(-1, 3)[-nf.value:])


SHADOW_CASES = ((5, 24, 50), (3, 24, 50))     # icosphere level, time steps, MC samples (CLI defaults)


def shadow_run(lib) -> dict:
    """The MC shadow flux (shadow_maker.c run_raytracer's computation) on the
    device for bumpy bodies of 20480 and 1280 faces at the CLI's defaults
    (za 12.3 deg, 24 time steps, 50 samples per face); wall time of
    dsmc_shadow_flux (normals, BVH build, upload, kernel, result back)."""
    out = {}
    for lev, tt, nm in SHADOW_CASES:
        v, f = bumpy_mesh(lib, lev)
        F = np.zeros(len(f) * tt)
        ms = []
        for _ in range(3):
            t0 = time.perf_counter()
            if lib.dsmc_shadow_flux(len(v), ptr(v, C.c_double), len(f), ptr(f, C.c_int32), C.c_double(12.3), tt, nm,
                                    ptr(F, C.c_double)):
                raise RuntimeError("dsmc_shadow_flux failed")
            ms.append((time.perf_counter() - t0) * 1e3)
        rays = int(np.count_nonzero(F.reshape(len(f), tt) > 0) * nm)
        out[f"{len(f)}_faces"] = {"ms": round(min(ms), 3), "lit_face_steps": int(np.count_nonzero(F > 0)),
                                  "rays_per_s": rays / (min(ms) / 1e3)}
    return out


def shadow_cpu(lib, threads: int) -> dict:
    """The oracle's shadow flux (occlusion tested against every face) on the
    host cores for the 1280-face body."""
    from tests.helpers import oracle
    orc = oracle(devmath=False)
    lev, tt, nm = SHADOW_CASES[1]
    v, f = bumpy_mesh(lib, lev)
    F = np.zeros(len(f) * tt)
    t0 = time.perf_counter()
    orc.cpu_ref_shadow_flux(len(v), ptr(v, C.c_double), len(f), ptr(f, C.c_int32), C.c_double(12.3), tt, nm,
                            ptr(F, C.c_double), threads)
    return {"cores": threads, f"{len(f)}_faces_ms": round((time.perf_counter() - t0) * 1e3, 3)}


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
    kms, fms, lean = [], [], []
    rounds, fl = C.c_int(0), C.c_double(0)
    lms, ln, lst = C.c_double(0), C.c_int(0), C.c_int64(0)
    lib.dsmc_run_sync(run)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if lib.dsmc_run_launch(run, shape_p) or lib.dsmc_run_sync(run):
            raise RuntimeError("timed launch failed")
        kms.append(lib.dsmc_run_kernel_ms(run))
        lib.dsmc_run_timing(run, C.byref(fl), C.byref(rounds))
        fms.append(fl.value)
        lib.dsmc_run_lean_timing(run, C.byref(lms), C.byref(ln), C.byref(lst))
        lean.append((lms.value, ln.value, lst.value))
    lib.dsmc_run_sync(run)
    wall = time.perf_counter() - t0
    if lib.dsmc_run_collect(run, C.byref(outs.result)):
        raise RuntimeError("collect failed")
    lib.dsmc_run_destroy(run)
    steps_per_run = outs.result.total_steps
    favg = float(np.mean(fms)) / 1e3          # flight-kernel device time per run
    launches = max(rounds.value, 1)           # flight launches per run
    traffic = PMC_TRAFFIC_1E7 if (shape is None and n == 10_000_000 and not args.no_moments) else None
    # compute-bound FP64 on the vector ALUs (no matrix core is used): the peak is
    # MI355X's FP64 vector rate, 78.6 TFLOP/s.  The roofline is the dominant
    # kernel's: the event-batched flight kernel when the run takes it (its own
    # launches, HIP-event time and device-counted particle-steps), else the
    # flight kernel with the landing service inline (every flight launch).
    lean_ms = sum(x[0] for x in lean) / len(lean) if lean else 0.0
    lean_n = lean[-1][1] if lean else 0
    lean_steps = lean[-1][2] if lean else 0
    if lean_n > 0 and lean_ms > 0:
        kname, k_launches, k_ms, k_steps = "lean_flight_kernel", lean_n, lean_ms, lean_steps
    else:
        kname = "phys_kernel<false>" if (shape is None or shape.service_inline) else "lean_flight_kernel"
        k_launches, k_ms, k_steps = launches, favg * 1e3, steps_per_run
    achieved = FLOPS_PER_STEP * k_steps / (k_ms / 1e3) / 1e12
    return {"params": p, "value": steps_per_run * args.steps / wall, "ms_per_run": wall / args.steps * 1e3,
            "particles": n, "particle_steps_per_run": steps_per_run,
            "fates": {k: int(getattr(outs.result, k)) for k in
                      ("n_escaped", "n_frozen", "n_photolysed", "n_capped")},
            "run_device_ms": [round(x, 3) for x in kms],
            "flight_kernel_ms": [round(x, 3) for x in fms], "rounds": rounds.value,
            "flight_all": {"ms_per_run": favg * 1e3, "launches_per_run": launches,
                           "achieved_tflops": FLOPS_PER_STEP * steps_per_run / favg / 1e12,
                           "note": "every flight launch of a run (the dominant kernel and the late launch "
                                   "with the landing service inline)"},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                         "compute": "FP64 vector ALU (VALU), not matrix cores",
                         "flops_per_particle_step": FLOPS_PER_STEP,
                         "flops_per_launch": FLOPS_PER_STEP * k_steps / k_launches,
                         "launches_per_run": k_launches,
                         "particle_steps_per_run_in_kernel": k_steps,
                         "kernel": kname, "kernel_avg_ms_per_launch": k_ms / k_launches,
                         "kernel_ms_per_run": k_ms}}


def progress(msg: str) -> None:
    """One progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def percentiles(x) -> dict:
    x = np.asarray(x, dtype=np.float64)
    if x.size == 0:
        return {"p50_ms": None, "p99_ms": None, "max_ms": None}
    return {"p50_ms": float(np.percentile(x, 50)), "p99_ms": float(np.percentile(x, 99)),
            "max_ms": float(x.max())}


def batch_once(lib, args, shape, params, surfs, moments: bool) -> dict:
    """Creates (and autotunes) the batch, launches it args.warmup + args.steps times."""
    pa, sa = W.as_arrays(params, surfs)
    n = len(params)
    st = C.c_int(0)
    flags = 0 if moments else capi.BATCH_DENSITY_ONLY
    t_tune = time.perf_counter()
    b = lib.dsmc_batch_create(n, pa, sa, flags, C.byref(shape) if shape is not None else None, C.byref(st))
    if not b:
        raise RuntimeError(f"dsmc_batch_create failed: {st.value}")
    t_tune = time.perf_counter() - t_tune
    bufs = [Buffers(p, 5, per_particle=False, moments=moments) for p in params]
    arr = (C.POINTER(capi.Result) * n)(*[C.pointer(x.result) for x in bufs])
    warm, reps = min(args.warmup, 1), max(1, min(args.steps, 3))     # each launch is a whole batch
    for _ in range(warm):
        if lib.dsmc_batch_launch(b, arr):
            raise RuntimeError("batch warmup failed")
    done = np.zeros(n)
    t0 = time.perf_counter()
    for _ in range(reps):
        if lib.dsmc_batch_launch(b, arr):
            raise RuntimeError("batch launch failed")
    wall = time.perf_counter() - t0
    w, rounds, used = C.c_double(), C.c_int64(), capi.LaunchShape()
    lib.dsmc_batch_timing(b, C.byref(w), capi.ptr(done, C.c_double), C.byref(rounds), C.byref(used))
    lib.dsmc_batch_destroy(b)
    steps = sum(int(x.result.total_steps) for x in bufs)
    return {"value": steps * reps / wall, "particle_steps_per_batch": steps, "launches_timed": reps,
            "ms_per_batch": wall / reps * 1e3, "rounds": int(rounds.value),
            "run_finish_ms": percentiles(done), "autotune_s": round(t_tune, 3),
            "shape": {"block": used.block, "blocks_per_cu": used.blocks_per_cu,
                      "relaunch_batch": used.relaunch_batch, "step_budget": used.step_budget,
                      "service_inline": used.service_inline}}


def batched_run(lib, args, shape) -> dict:
    """The batched workload: 64 independent runs (counts log-uniform on [1e4, 2e6],
    seed 1; body and subsolar longitude drawn per run), icosphere 2562/5120, a
    40x36x72 FP64 density grid per run, in one pool; the same batch with velocity
    moments per cell too; and the same runs one after another through the plain
    single-run path (density grid, default launch shape)."""
    specs = W.batch_specs(args.batch_runs, seed=1, scale=args.batch_scale)
    params, surfs = W.build_runs(lib, specs, levels=4, hd=40)
    main = batch_once(lib, args, shape, params, surfs, moments=False)
    out = {"value": main.pop("value"), "unit": "particle-steps/s", "runs": len(specs),
           "particles": int(sum(s.n for s in specs)), "outputs": "density grid per run", **main,
           "graph": "hipGraph of 8 rounds (flight [+service] + reset) replayed"}
    if args.batch_moments:
        wm = batch_once(lib, args, shape, params, surfs, moments=True)
        out["with_velocity_moments"] = {"value": wm["value"], "ms_per_batch": wm["ms_per_batch"],
                                        "shape": wm["shape"], "autotune_s": wm["autotune_s"]}
    if args.sequential:
        st = C.c_int(0)
        t_seq, seq_steps = 0.0, 0
        for p, s in zip(params, surfs):
            ob = Buffers(p, 5, per_particle=False, moments=False)
            run = lib.dsmc_run_create(C.byref(p), C.byref(s.surface), C.byref(ob.result), C.byref(st))
            if not run:
                raise RuntimeError(f"dsmc_run_create failed: {st.value}")
            t1 = time.perf_counter()
            if lib.dsmc_run_launch(run, None) or lib.dsmc_run_sync(run):
                raise RuntimeError("plain launch failed")
            t_seq += time.perf_counter() - t1
            lib.dsmc_run_collect(run, C.byref(ob.result))
            seq_steps += int(ob.result.total_steps)
            lib.dsmc_run_destroy(run)
        out["plain_sequential"] = {"value": seq_steps / t_seq, "ms": t_seq * 1e3, "particle_steps": seq_steps}
        out["speedup_vs_plain_sequential"] = out["value"] / (seq_steps / t_seq)
    return out


def collisional_run(lib, args) -> dict:
    """The collisional workload: Ceres H2O, phase 1 (collisionless, speed
    histograms) then phase 2 (hard spheres d = 4.6e-10 m, dt 5 s, degree-6
    partner fits), emission 2e26 molecules/s x10, grid 50x36x72, both phases at
    args.coll_particles, through the device-resident two-phase pipeline."""
    surf = icosphere(lib, 3)
    n = args.coll_particles
    p2 = capi.preset(lib, capi.PRESET_CERES_PHYSICAL_COLL, 36, 72, n=n, seed=args.seed)
    p1 = capi.preset(lib, capi.PRESET_CERES_PHYSICAL, 36, 72, n=n, seed=args.seed)
    p1.production_rate = p2.production_rate      # ~6 kg/s of water ...
    p1.sublimation_scale = p2.sublimation_scale  # ... x10
    r1 = Buffers(p1, 5, per_particle=False, moments=True, vhist=True)
    r2 = Buffers(p2, 5, per_particle=False, moments=True)
    t0 = time.perf_counter()
    rc = lib.dsmc_run_collisional(C.byref(p1), C.byref(p2), C.byref(surf.surface), C.byref(r1.result),
                                  C.byref(r2.result))
    wall = time.perf_counter() - t0
    if rc:
        raise RuntimeError(f"dsmc_run_collisional: {rc}")
    s1, s2 = int(r1.result.total_steps), int(r2.result.total_steps)
    return {"value": (s1 + s2) / wall, "unit": "particle-steps/s", "particles_per_phase": n,
            "wall_ms": wall * 1e3, "phase1_steps": s1, "phase2_steps": s2,
            "collisions": int(r2.result.n_collisions), "collisions_per_phase2_step": r2.result.n_collisions / max(s2, 1),
            "phase2_fates": {k: int(getattr(r2.result, k)) for k in
                             ("n_escaped", "n_frozen", "n_photolysed", "n_capped")}}


def served_run(lib, args, shape, rate: float, seconds: float, drain: bool = True) -> dict:
    """Poisson request stream (`rate` req/s for `seconds`) of runs, counts
    log-uniform on [1e4, 4e6], body and subsolar longitude drawn per request, 30%
    collisional (phase 1 + phase 2 at the request's count), each returning its
    40x36x72 density grid, through the continuous-batching server; latency =
    arrival to result ready.  drain: serve until every request completed;
    otherwise stop when the arrival window closes and report what completed in
    it, the queue's growth, and latencies with the unfinished requests counted at
    their age when the window closed (a lower bound)."""
    stream = W.request_stream(rate, seconds, seed=5, scale=args.serve_scale,
                              frac_collisional=args.serve_collisional)
    specs = [s for _, s in stream]
    params, surfs = W.build_runs(lib, specs, levels=4, hd=40)
    p2s = [None] * len(specs)
    for i, sp in enumerate(specs):
        if sp.collisional:
            params[i], p2s[i] = W.coll_params(lib, sp.body, sp.n, sp.lon, sp.seed, hd=40)
    sv = lib.dsmc_server_create(args.serve_budget_gb << 30, args.serve_max_inflight, 2)  # collisional accepted
    if not sv:
        raise RuntimeError("dsmc_server_create failed")
    bufs = [Buffers(p2s[i] if p2s[i] is not None else p, 5, per_particle=False, moments=False)
            for i, p in enumerate(params)]
    tickets, t_sub = [0] * len(stream), [0.0] * len(stream)
    t0 = time.perf_counter()
    i, outstanding = 0, set()
    first = True
    t_report = 0.0
    growth = []
    while i < len(stream) or outstanding:
        if time.perf_counter() - t0 > t_report + 5.0:
            t_report = time.perf_counter() - t0
            growth.append({"t_s": round(t_report, 1), "submitted": i, "outstanding": len(outstanding),
                           "completed": i - len(outstanding)})
            progress(f"served: {i}/{len(stream)} submitted, {len(outstanding)} outstanding at {t_report:.0f} s")
        now = time.perf_counter() - t0
        if not drain and now >= seconds:
            break
        while i < len(stream) and stream[i][0] <= now:
            tk = C.c_uint64()
            if p2s[i] is not None:
                rc = lib.dsmc_server_submit_collisional(sv, C.byref(params[i]), C.byref(p2s[i]),
                                                        C.byref(surfs[i].surface), C.byref(bufs[i].result),
                                                        C.byref(tk))
            else:
                rc = lib.dsmc_server_submit(sv, C.byref(params[i]), C.byref(surfs[i].surface),
                                            C.byref(bufs[i].result), C.byref(tk))
            if rc:
                raise RuntimeError(f"submit failed: {rc}")
            if first:
                if shape is not None and lib.dsmc_server_set_shape(sv, C.byref(shape)):
                    raise RuntimeError("set_shape failed")
                first = False
            tickets[i], t_sub[i] = tk.value, time.perf_counter() - t0
            outstanding.add(i)
            i += 1
        if outstanding:
            if lib.dsmc_server_step(sv) < 0:
                raise RuntimeError("server step failed")
            for j in list(outstanding):
                st = lib.dsmc_server_poll(sv, tickets[j])
                if st < 0:
                    raise RuntimeError(f"request {j} failed: {st}")
                if st == 1:
                    outstanding.discard(j)
        elif i < len(stream):
            time.sleep(max(0.0, stream[i][0] - (time.perf_counter() - t0)))
    t_end = time.perf_counter() - t0
    stats = capi.ServerStats()
    lib.dsmc_server_get_stats(sv, C.byref(stats))
    unfinished = sorted(outstanding)
    lib.dsmc_server_destroy(sv)         # waits for the replay in flight; unfinished requests are dropped
    done = [j for j in range(i) if j not in outstanding]
    lat = [bufs[j].result.elapsed_ms + (t_sub[j] - stream[j][0]) * 1e3 for j in done]
    steps = int(stats.particle_steps)
    ncoll = sum(1 for s in specs if s.collisional)
    out = {"value": steps / t_end, "unit": "particle-steps/s", "requests": len(stream),
           "collisional_requests": ncoll, "offered_seconds": seconds, "rate_per_s": rate, "wall_s": t_end,
           "particles": int(sum(s.n for s in specs)), "particle_steps": steps, "latency": percentiles(lat),
           "rounds": int(stats.rounds), "slots": int(stats.slots), "queue_growth": growth}
    if drain:
        # offered load: the stream's whole work over the arrival window; above the
        # served rate the queue grows for the whole window and latency with it
        out["offered_steps_per_s"] = steps / seconds
    else:
        # the work of the requests completed in the window, their latency, and
        # every request's latency with the unfinished ones at their age at the
        # window's end (a lower bound on what they would take)
        ages = lat + [(t_end - stream[j][0]) * 1e3 for j in unfinished] + \
               [(t_end - stream[j][0]) * 1e3 for j in range(i, len(stream)) if stream[j][0] <= t_end]
        out.update({"completed": len(done), "unfinished_at_window_end": len(unfinished),
                    "latency_completed": out.pop("latency"),
                    "latency_all_lower_bound": percentiles(ages),
                    "value_note": "particle-steps of the requests completed in the window / window length; "
                                  "work done on unfinished requests is not counted"})
    return out


def dropin_run(lib, args) -> dict:
    """The work behind one drop-in run_ballistic call at the reference's own size
    (compat ballistic, 1e4 particles, 36x72 surface field, 100 uniform shells):
    the device run (staging, particles, results back) and its four text files,
    written and read back (as run_advanced reads them) with the default host
    threads and with one."""
    import tempfile
    ftd, ffd, n = 36, 72, 10_000
    lat = (np.arange(ftd) + 0.5) / ftd * np.pi - np.pi / 2
    lon = (np.arange(ffd) + 0.5) / ffd * 2 * np.pi
    T = 80.0 + 160.0 * np.maximum(np.outer(np.cos(lat), np.sin(lon)), 0.0) ** 0.25
    T = np.ascontiguousarray(T.ravel())
    p = capi.preset(lib, PRESET_COMPAT_BALLISTIC, ftd, ffd, n=n, seed=args.seed)
    surf = SurfaceArrays(T, ftd, ffd)
    b = Buffers(p, 5, per_particle=True, vrec=True)
    dev = []
    for _ in range(3):
        t0 = time.perf_counter()
        if lib.dsmc_run_ballistic(C.byref(p), C.byref(surf.surface), C.byref(b.result)):
            raise RuntimeError("dsmc_run_ballistic failed")
        dev.append((time.perf_counter() - t0) * 1e3)
    nrec = int(b.result.vrec_n)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    lib.read_density_dat.restype = dp
    lib.read_v_record_dat.argtypes = [C.c_char_p, C.POINTER(dp), C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), ip]
    out = {"particles": n, "particle_steps": int(b.result.total_steps), "v_records": nrec,
           "device_ms": round(float(np.median(dev)), 2)}
    with tempfile.TemporaryDirectory() as d:
        for label, threads in (("", 0), ("_1thread", 1)):
            lib.dsmc_io_threads(threads)
            f = {k: os.path.join(d, f"r{label}_{k}.dat").encode() for k in ("den", "horizon", "v_record", "ttime")}
            w, r = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                rc = (lib.write_density_dat(f["den"], b.density.ctypes.data_as(dp), ftd, ffd, p.hd) |
                      lib.write_horizon_dat(f["horizon"], b.horizon.ctypes.data_as(dp), ftd, ffd) |
                      lib.write_v_record_dat(f["v_record"], b.vrec.ctypes.data_as(dp), b.vrec_td.ctypes.data_as(ip),
                                             b.vrec_fd.ctypes.data_as(ip), b.vrec_ht.ctypes.data_as(ip), nrec) |
                      lib.write_ttime_dat(f["ttime"], b.ttime.ctypes.data_as(dp), 5, n))
                w.append((time.perf_counter() - t0) * 1e3)
                if rc:
                    raise RuntimeError("writing the .dat files failed")
                a1, a2, a3 = C.c_int(), C.c_int(), C.c_int()
                v, vt, vf, vh, nr = dp(), ip(), ip(), ip(), C.c_int()
                t0 = time.perf_counter()
                den = lib.read_density_dat(f["den"], C.byref(a1), C.byref(a2), C.byref(a3), 1)
                ok = lib.read_v_record_dat(f["v_record"], C.byref(v), C.byref(vt), C.byref(vf), C.byref(vh),
                                           C.byref(nr)) == 0
                r.append((time.perf_counter() - t0) * 1e3)
                if not den or not ok or nr.value != nrec:
                    raise RuntimeError("reading the .dat files back failed")
            out[f"write_ms{label}"] = round(min(w), 2)
            out[f"read_ms{label}"] = round(min(r), 2)
        out["text_bytes"] = sum(os.path.getsize(x) for x in f.values())
    lib.dsmc_io_threads(0)
    out["io_threads"] = min(16, os.cpu_count() or 1)
    return out


def copy_calibration(mib: int) -> None:
    """A plain device-to-device copy of known size: what the HBM byte counters
    read for it (per dispatch) against its 2 x size algorithmic bytes gives the
    factor that scales the counters of the particle kernels."""
    import torch
    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 5}))


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
    ap.add_argument("--step-budget", type=int, default=0)
    ap.add_argument("--service-inline", type=int, default=0)
    ap.add_argument("--late-step-budget", type=int, default=0, help="explicit shape: budget after two rounds")
    ap.add_argument("--park-batch", type=int, default=0, help="explicit shape: parked lanes per cell-logic pass")
    ap.add_argument("--late-inline", type=int, default=0,
                    help="explicit shape: rounds after the first two in one launch with the service inline")
    ap.add_argument("--cpu-sample", type=int, default=1500000,
                    help="particles of the all-core CPU baseline sample (~15 s on 16 cores)")
    ap.add_argument("--cpu-single-sample", type=int, default=60000,
                    help="particles of the one-thread CPU baseline sample (~10 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-moments", action="store_true", help="density only (no velocity moments)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-runs", type=int, default=64)
    ap.add_argument("--batch-scale", type=float, default=1.0, help="multiplies every run's particle count")
    ap.add_argument("--sequential", type=int, default=1, help="also time the batch's runs one by one")
    ap.add_argument("--batch-moments", type=int, default=1, help="also time the batch with velocity moments")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-seconds", type=float, default=120.0)
    ap.add_argument("--serve-sustained-rate", type=float, default=2.0, help="second served leg (0: none)")
    ap.add_argument("--serve-sustained-seconds", type=float, default=10.0)
    ap.add_argument("--serve-scale", type=float, default=1.0)
    ap.add_argument("--serve-budget-gb", type=int, default=128)
    ap.add_argument("--serve-max-inflight", type=int, default=256, help="requests the server holds at once")
    ap.add_argument("--serve-collisional", type=float, default=0.3, help="fraction of collisional requests")
    ap.add_argument("--no-coll", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in run_ballistic files leg")
    ap.add_argument("--no-thermal", action="store_true", help="skip the thermal model leg")
    ap.add_argument("--no-shadow", action="store_true", help="skip the MC shadow flux leg")
    ap.add_argument("--coll-particles", type=int, default=2_000_000)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy a buffer of this many MiB device to device 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the path is not sharded across GPUs)"}))
        sys.exit(2)
    if args.calibrate_copy_mib:
        copy_calibration(args.calibrate_copy_mib)
        return

    lib = capi.load()
    surf = icosphere(lib, 3)
    shape = None
    if (args.block or args.blocks_per_cu or args.relaunch_batch or args.step_budget or args.service_inline or
            args.late_step_budget or args.park_batch or args.late_inline):
        shape = capi.LaunchShape(args.block or 256, args.blocks_per_cu or 4, args.relaunch_batch or 16,
                                 args.step_budget or 1024, args.service_inline, args.late_step_budget,
                                 args.park_batch, args.late_inline)
    progress(f"plain run at {args.particles:.0e} particles")
    main_run = plain_run(lib, surf, args.particles, args, shape)
    progress(f"plain: {main_run['value']:.3e} particle-steps/s")
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
        progress(f"plain run at {args.also_particles:.0e} particles")
        other = plain_run(lib, surf, args.also_particles, args, shape)
        other.pop("params")
        line["config"][f"plain_{args.also_particles:.0e}"] = {
            "value": other["value"], "ms_per_run": other["ms_per_run"],
            "flight_kernel_ms": other["flight_kernel_ms"], "rounds": other["rounds"]}
    if not args.no_dropin:
        progress("drop-in run_ballistic work at 1e4 particles (device run + .dat files)")
        line["dropin_files"] = dropin_run(lib, args)
    if not args.no_thermal:
        progress("thermal model (36x72, 360x720)")
        line["thermal"] = thermal_run(lib)
    if not args.no_shadow:
        progress("MC shadow flux (20480 and 1280 faces)")
        line["shadow"] = shadow_run(lib)
    tuned = None
    if not args.no_batch:
        progress("batched: 64 runs (autotune, launches, plain sequential)")
        bt = batched_run(lib, args, shape)
        progress(f"batched: {bt['value']:.3e} particle-steps/s")
        line["batched"] = bt
        s = bt["shape"]
        tuned = capi.LaunchShape(s["block"], s["blocks_per_cu"], s["relaunch_batch"], s["step_budget"],
                                 s["service_inline"])
        line["batched_vs_plain"] = bt["value"] / main_run["value"]
    if not args.no_coll:
        progress(f"collisional pipeline at {args.coll_particles:.0e} particles per phase")
        line["collisional"] = collisional_run(lib, args)
        progress(f"collisional: {line['collisional']['value']:.3e} particle-steps/s")
    if not args.no_serve:
        sshape = shape if shape is not None else tuned
        progress(f"served (Poisson stream): {args.serve_rate} req/s for {args.serve_seconds} s")
        line["served_cfg5"] = served_run(lib, args, sshape, args.serve_rate, args.serve_seconds, drain=False)
        progress(f"served: {line['served_cfg5']['value']:.3e} particle-steps/s, "
                 f"{line['served_cfg5']['completed']}/{line['served_cfg5']['requests']} completed in the window")
        if args.serve_sustained_rate > 0:
            # the same stream at a rate the device keeps up with: latency is
            # then service time, not time spent queued behind a backlog
            progress(f"served (sustainable load): {args.serve_sustained_rate} req/s for "
                     f"{args.serve_sustained_seconds} s")
            line["served_sustained"] = served_run(lib, args, sshape, args.serve_sustained_rate,
                                                  args.serve_sustained_seconds)
            progress(f"served (sustainable load): p50 {line['served_sustained']['latency']['p50_ms']:.0f} ms")
    if not args.no_cpu:
        progress("CPU baseline")
        host = host_cpus()
        threads = cpu_threads(host)
        line["cpu_baseline"] = cpu_baseline(p, surf, args.cpu_sample, threads, args.cpu_single_sample)
        line["cpu_baseline"]["host"] = host
        if not args.no_thermal:
            line["cpu_baseline"]["thermal"] = thermal_cpu(threads)
        if not args.no_shadow:
            line["cpu_baseline"]["shadow"] = shadow_cpu(lib, threads)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
