"""Multi-run pools on the GPU: the batch API and the serving loop.

Each run of a batch, and each served request, must give the same results as the
same run alone: per-particle end states, step counts, fates and hop counts bit
for bit, integer deposit counts, horizon (sums of unit weights) and fate
counters exactly, and velocity moments to the summation-order tolerance
|delta| <= 1e-12 * count * 5e3 m/s + 1e-9 (every term is one particle velocity
below 5 km/s; only the order of the FP64 atomic additions differs).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from dsmc_suite_amd import capi
from dsmc_suite_amd import workloads as W
from dsmc_suite_amd.capi import Buffers

from tests.helpers import oracle

pytestmark = pytest.mark.gpu

LEVELS, HD = 3, 40
COUNTERS = ("n_escaped", "n_frozen", "n_photolysed", "n_depleted", "n_capped", "n_collisions", "total_steps")


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def specs():
    return [W.RunSpec(W.BODIES[0], 3000, 0.3, 11), W.RunSpec(W.BODIES[1], 700, 2.0, 12),
            W.RunSpec(W.BODIES[2], 2500, 4.1, 13), W.RunSpec(W.BODIES[3], 1500, 5.5, 14),
            W.RunSpec(W.BODIES[0], 257, 1.0, 15)]


def alone(lib, p, surf):
    b = Buffers(p, 5, per_particle=True, moments=True)
    st = C.c_int(0)
    run = lib.dsmc_run_create(C.byref(p), C.byref(surf.surface), C.byref(b.result), C.byref(st))
    assert run, st.value
    assert lib.dsmc_run_launch(run, None) == 0
    assert lib.dsmc_run_sync(run) == 0
    assert lib.dsmc_run_collect(run, C.byref(b.result)) == 0
    lib.dsmc_run_destroy(run)
    return b, {k: int(getattr(b.result, k)) for k in COUNTERS}


def assert_same_run(a: Buffers, ca: dict, b: Buffers, cb: dict, what: str):
    assert ca == cb, what
    for f in ("end_state", "steps", "fate", "hops", "vcount", "density", "horizon"):
        x, y = getattr(a, f), getattr(b, f)
        assert np.array_equal(x, y), f"{what}: {f} differs at {np.flatnonzero(x != y)[:5]}"
    cnt = np.repeat(a.vcount.astype(np.float64), 3)
    tol = 1e-12 * cnt * 5e3 + 1e-9
    d = np.abs(a.vmom - b.vmom)
    assert np.all(d <= tol), f"{what}: vmom max excess {np.max(d - tol)}"


def test_batch_runs_equal_single_runs(lib):
    sp = specs()
    params, surfs = W.build_runs(lib, sp, levels=LEVELS, hd=HD)
    bufs = [Buffers(p, 5, per_particle=True, moments=True) for p in params]
    pa, sa = W.as_arrays(params, surfs)
    res = (capi.Result * len(bufs))(*[b.result for b in bufs])
    shape = capi.LaunchShape(256, 4, 16, 1024, 1)
    print(f"batch of {len(sp)} runs, counts {[s.n for s in sp]}", flush=True)
    assert lib.dsmc_run_batch(len(sp), pa, sa, res, C.byref(shape)) == 0
    for i, (p, s) in enumerate(zip(params, surfs)):
        b1, c1 = alone(lib, p, s)
        cb = {k: int(getattr(res[i], k)) for k in COUNTERS}
        assert cb["total_steps"] > 0
        assert sum(cb[k] for k in COUNTERS[:5]) == p.nparticles
        assert_same_run(b1, c1, bufs[i], cb, f"run {i} ({sp[i].body.name}, n={sp[i].n})")


def oracle_ballistic(orc, p, surf):
    ob = Buffers(p, 5, per_particle=True, moments=True)
    assert orc.cpu_ref_ballistic(C.byref(p), C.byref(surf.surface), C.byref(ob.result), 8, 0) == 0
    return ob, {k: int(getattr(ob.result, k)) for k in COUNTERS}


def oracle_collisional(orc, p1, p2, surf):
    orc.cpu_ref_collisional.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_int]
    w1 = Buffers(p1, 5, per_particle=False, moments=True, vhist=True)
    w2 = Buffers(p2, 5, per_particle=True, moments=True)
    assert orc.cpu_ref_collisional(C.byref(p1), C.byref(p2), C.byref(surf.surface), C.byref(w1.result),
                                   C.byref(w2.result), 8, 0) == 0
    return w2, {k: int(getattr(w2.result, k)) for k in COUNTERS}


def assert_matches_oracle(got: Buffers, cg: dict, want: Buffers, cw: dict, what: str):
    """HIP run against the oracle's device-math build: per-particle results,
    deposit counts and fate counters bit for bit, velocity moments to the
    summation-order tolerance."""
    assert cg == cw, what
    for f in ("end_state", "steps", "fate", "hops", "vcount"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f"{what}: {f}"
    cnt = np.repeat(want.vcount.astype(np.float64), 3)
    assert np.all(np.abs(got.vmom - want.vmom) <= 1e-12 * cnt * 5e3 + 1e-9), what


@pytest.mark.parametrize("levels", [3, 4])
def test_batch_run_matches_oracle(lib, levels):
    """Two runs of an autotuned batch (Ceres and the Moon) against the CPU oracle
    (device-math build), on the 1280-face icosphere and on the batched
    workload's 5120-face one (its 64 x 64 cube-map face table)."""
    sp = specs()[:2]
    params, surfs = W.build_runs(lib, sp, levels=levels, hd=HD)
    bufs = [Buffers(p, 5, per_particle=True, moments=True) for p in params]
    pa, sa = W.as_arrays(params, surfs)
    res = (capi.Result * 2)(*[b.result for b in bufs])
    print(f"batch vs oracle, icosphere levels {levels}", flush=True)
    assert lib.dsmc_run_batch(2, pa, sa, res, None) == 0     # autotuned shape
    orc = oracle(devmath=True)
    for i in range(2):
        ob, co = oracle_ballistic(orc, params[i], surfs[i])
        assert_matches_oracle(bufs[i], {k: int(getattr(res[i], k)) for k in COUNTERS}, ob, co,
                              f"run {i} ({sp[i].body.name}, {levels} levels)")


@pytest.mark.parametrize("levels", [3, 4])
def test_served_requests_match_oracle(lib, levels):
    """Requests served together - two ballistic (the Moon, Mercury) and one
    collisional (Ceres: phase 1, fits and number density on the device, phase 2)
    - each against the CPU oracle (device-math build)."""
    sp = [W.RunSpec(W.BODIES[1], 700, 2.0, 31), W.RunSpec(W.BODIES[3], 1200, 5.0, 32)]
    params, surfs = W.build_runs(lib, sp, levels=levels, hd=HD)
    csurf = W.icosphere(lib, levels, W.BODIES[0], 0.4)
    p1, p2 = W.coll_params(lib, W.BODIES[0], 1500, 0.4, 33, hd=HD)
    p2.nparticles = 600
    sv = lib.dsmc_server_create(4 << 30, 8, 3)          # per-particle outputs, collisional requests
    assert sv
    try:
        outs = [Buffers(p, 5, per_particle=True, moments=True) for p in params]
        cout = Buffers(p2, 5, per_particle=True, moments=True)
        tk = []
        for j in range(2):
            t = C.c_uint64()
            assert lib.dsmc_server_submit(sv, C.byref(params[j]), C.byref(surfs[j].surface), C.byref(outs[j].result),
                                          C.byref(t)) == 0
            tk.append(t.value)
            if j == 0:
                assert lib.dsmc_server_submit_collisional(sv, C.byref(p1), C.byref(p2), C.byref(csurf.surface),
                                                          C.byref(cout.result), C.byref(t)) == 0
                tk.append(t.value)
            assert lib.dsmc_server_step(sv) >= 0
        print(f"served requests vs oracle, icosphere levels {levels}", flush=True)
        for x in tk:
            assert lib.dsmc_server_wait(sv, x) == 0
    finally:
        lib.dsmc_server_destroy(sv)
    orc = oracle(devmath=True)
    for j in range(2):
        ob, co = oracle_ballistic(orc, params[j], surfs[j])
        assert_matches_oracle(outs[j], {k: int(getattr(outs[j].result, k)) for k in COUNTERS}, ob, co,
                              f"served ballistic request {j} ({sp[j].body.name})")
    w2, cw = oracle_collisional(orc, p1, p2, csurf)
    assert cw["n_collisions"] > 10, cw
    assert_matches_oracle(cout, {k: int(getattr(cout.result, k)) for k in COUNTERS}, w2, cw,
                          "served collisional request")


def test_staged_batch_relaunch(lib):
    """A staged batch launched twice gives the same results; timing is reported."""
    sp = specs()[:3]
    params, surfs = W.build_runs(lib, sp, levels=LEVELS, hd=HD)
    pa, sa = W.as_arrays(params, surfs)
    st = C.c_int(0)
    b = lib.dsmc_batch_create(3, pa, sa, 0, None, C.byref(st))
    assert b, st.value
    outs = []
    for _ in range(2):
        bufs = [Buffers(p, 5, per_particle=False, moments=True) for p in params]
        arr = (C.POINTER(capi.Result) * 3)(*[C.pointer(x.result) for x in bufs])
        assert lib.dsmc_batch_launch(b, arr) == 0
        outs.append(bufs)
    wall, rounds = C.c_double(), C.c_int64()
    done = np.zeros(3)
    shape = capi.LaunchShape()
    assert lib.dsmc_batch_timing(b, C.byref(wall), capi.ptr(done, C.c_double), C.byref(rounds), C.byref(shape)) == 0
    lib.dsmc_batch_destroy(b)
    assert wall.value > 0 and rounds.value > 0 and np.all(done > 0) and np.all(done <= wall.value + 1e-6)
    assert shape.block in (64, 128, 192, 256)
    for x, y in zip(*outs):
        assert np.array_equal(x.vcount, y.vcount) and np.array_equal(x.horizon, y.horizon)
        assert x.result.total_steps == y.result.total_steps


def test_batch_density_only_equals_full(lib):
    """A density-only batch (one deposit atomic per cell visit) gives the full
    batch's per-particle results, counts, horizon and fate counters bit for bit;
    velocity moments it was not asked for come back zero."""
    sp = specs()
    params, surfs = W.build_runs(lib, sp, levels=LEVELS, hd=HD)
    pa, sa = W.as_arrays(params, surfs)
    shape = capi.LaunchShape(256, 4, 16, 1024, 1)
    st = C.c_int(0)
    outs = []
    for flags in (capi.BATCH_PER_PARTICLE, capi.BATCH_PER_PARTICLE | capi.BATCH_DENSITY_ONLY):
        print(f"batch flags {flags}", flush=True)
        b = lib.dsmc_batch_create(len(sp), pa, sa, flags, C.byref(shape), C.byref(st))
        assert b, st.value
        bufs = [Buffers(p, 5, per_particle=True, moments=True) for p in params]
        for x in bufs:
            x.vmom[:] = 1.0                     # must be overwritten
        arr = (C.POINTER(capi.Result) * len(sp))(*[C.pointer(x.result) for x in bufs])
        assert lib.dsmc_batch_launch(b, arr) == 0
        lib.dsmc_batch_destroy(b)
        outs.append(bufs)
    for i, (x, y) in enumerate(zip(*outs)):
        for f in ("end_state", "steps", "fate", "hops", "vcount", "density", "horizon"):
            assert np.array_equal(getattr(x, f), getattr(y, f)), (i, f)
        assert {k: int(getattr(x.result, k)) for k in COUNTERS} == {k: int(getattr(y.result, k)) for k in COUNTERS}
        assert np.any(x.vmom != 0.0) and np.all(y.vmom == 0.0), i


def test_server_alone_equals_mix(lib):
    """Continuous batching: run A alone, then A admitted while other runs fly."""
    sp = specs()
    params, surfs = W.build_runs(lib, sp, levels=LEVELS, hd=HD)
    sv = lib.dsmc_server_create(2 << 30, 8, 1)
    assert sv
    try:
        t = C.c_uint64()
        ra = Buffers(params[0], 5, per_particle=True, moments=True)
        assert lib.dsmc_server_submit(sv, C.byref(params[0]), C.byref(surfs[0].surface), C.byref(ra.result),
                                      C.byref(t)) == 0
        assert lib.dsmc_server_wait(sv, t.value) == 0
        ca = {k: int(getattr(ra.result, k)) for k in COUNTERS}
        assert ra.result.elapsed_ms > 0

        mix = [Buffers(p, 5, per_particle=True, moments=True) for p in params]
        tickets = {}
        order = [1, 3, 0, 2, 4]
        saw_overlap = False
        for j, i in enumerate(order):
            tk = C.c_uint64()
            assert lib.dsmc_server_submit(sv, C.byref(params[i]), C.byref(surfs[i].surface),
                                          C.byref(mix[i].result), C.byref(tk)) == 0
            tickets[i] = tk.value
            assert lib.dsmc_server_step(sv) >= 0
            stt = capi.ServerStats()
            lib.dsmc_server_get_stats(sv, C.byref(stt))
            saw_overlap |= stt.running >= 2
        for i in order:
            assert lib.dsmc_server_wait(sv, tickets[i]) == 0
        assert saw_overlap, "requests never overlapped on the device"
        cm = {k: int(getattr(mix[0].result, k)) for k in COUNTERS}
        assert_same_run(ra, ca, mix[0], cm, "request A alone vs in the mix")
        for i in order[1:]:
            b1, c1 = alone(lib, params[i], surfs[i])
            ci = {k: int(getattr(mix[i].result, k)) for k in COUNTERS}
            assert_same_run(b1, c1, mix[i], ci, f"request {i}")
    finally:
        lib.dsmc_server_destroy(sv)


def test_server_budget_and_cancel(lib):
    sp = specs()
    params, surfs = W.build_runs(lib, sp, levels=LEVELS, hd=HD)
    sv = lib.dsmc_server_create(200 << 20, 4, 0)
    assert sv
    try:
        big = params[0].copy()
        big.nparticles = 10_000_000
        r = Buffers(big, 5, per_particle=False, moments=True)
        t = C.c_uint64()
        assert lib.dsmc_server_submit(sv, C.byref(big), C.byref(surfs[0].surface), C.byref(r.result),
                                      C.byref(t)) == -5     # DSMC_ERR_BUDGET
        ra = Buffers(params[0], 5, per_particle=False, moments=True)
        rb = Buffers(params[2], 5, per_particle=False, moments=True)
        ta, tb = C.c_uint64(), C.c_uint64()
        assert lib.dsmc_server_submit(sv, C.byref(params[0]), C.byref(surfs[0].surface), C.byref(ra.result),
                                      C.byref(ta)) == 0
        assert lib.dsmc_server_submit(sv, C.byref(params[2]), C.byref(surfs[2].surface), C.byref(rb.result),
                                      C.byref(tb)) == 0
        assert lib.dsmc_server_cancel(sv, tb.value) == 0
        assert lib.dsmc_server_poll(sv, tb.value) == -6     # DSMC_ERR_CANCELLED
        assert lib.dsmc_server_wait(sv, ta.value) == 0
        assert ra.result.total_steps > 0
        assert lib.dsmc_server_release(sv, ta.value) == 0
        assert lib.dsmc_server_poll(sv, ta.value) == -1
    finally:
        lib.dsmc_server_destroy(sv)


def test_server_collisional_alone_mix_and_pipeline(lib):
    """A collisional request served alone, served among ballistic requests, and
    run through dsmc_run_collisional: the same phase-2 particles bit for bit."""
    surf = W.icosphere(lib, LEVELS, W.BODIES[0], 0.0)
    p1, p2 = W.coll_params(lib, W.BODIES[0], 1500, 0.0, 71, hd=HD)
    p2.nparticles = 600
    ref = Buffers(p2, 5, per_particle=True, moments=True)
    assert lib.dsmc_run_collisional(C.byref(p1), C.byref(p2), C.byref(surf.surface), None, C.byref(ref.result)) == 0
    cref = {k: int(getattr(ref.result, k)) for k in COUNTERS}
    assert cref["n_collisions"] > 10, cref
    sp = specs()
    params, surfs = W.build_runs(lib, sp, levels=LEVELS, hd=HD)
    sv = lib.dsmc_server_create(4 << 30, 8, 3)          # per-particle outputs, collisional requests
    assert sv
    try:
        ra = Buffers(p2, 5, per_particle=True, moments=True)
        t = C.c_uint64()
        assert lib.dsmc_server_submit_collisional(sv, C.byref(p1), C.byref(p2), C.byref(surf.surface),
                                                  C.byref(ra.result), C.byref(t)) == 0
        assert lib.dsmc_server_wait(sv, t.value) == 0
        assert_same_run(ref, cref, ra, {k: int(getattr(ra.result, k)) for k in COUNTERS}, "served alone")
        mix = Buffers(p2, 5, per_particle=True, moments=True)
        others = [Buffers(p, 5, per_particle=True, moments=True) for p in params[:3]]
        tk = []
        for j in range(3):
            tt = C.c_uint64()
            assert lib.dsmc_server_submit(sv, C.byref(params[j]), C.byref(surfs[j].surface),
                                          C.byref(others[j].result), C.byref(tt)) == 0
            tk.append(tt.value)
            if j == 1:
                assert lib.dsmc_server_submit_collisional(sv, C.byref(p1), C.byref(p2), C.byref(surf.surface),
                                                          C.byref(mix.result), C.byref(t)) == 0
                tk.append(t.value)
            assert lib.dsmc_server_step(sv) >= 0
        for x in tk:
            assert lib.dsmc_server_wait(sv, x) == 0
        assert_same_run(ref, cref, mix, {k: int(getattr(mix.result, k)) for k in COUNTERS}, "served in a mix")
        for j in range(3):
            b1, c1 = alone(lib, params[j], surfs[j])
            assert_same_run(b1, c1, others[j], {k: int(getattr(others[j].result, k)) for k in COUNTERS},
                            f"ballistic request {j} beside a collisional one")
    finally:
        lib.dsmc_server_destroy(sv)

