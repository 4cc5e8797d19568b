"""dsmc_run_ballistic_noalloc against dsmc_run_ballistic, through the C ABI.

The allocation-free entry point takes every device buffer of a run from one
block of device memory that keeps the previous call's contents, so it is called
twice in a row, the second run smaller than the first, and each call's results
must equal dsmc_run_ballistic's on the same inputs: per-particle outputs, deposit
counts, fate counters and (compat) transit times bit for bit, floating-point
grid sums to the summation-order tolerance (atomics add in any order), velocity
records per (td, ht) cell by count.  A run too large for the block is rejected
with DSMC_ERR_BUDGET.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from dsmc_suite_amd import capi
from dsmc_suite_amd import workloads as W
from dsmc_suite_amd.capi import PRESET_CERES_PHYSICAL, Buffers, SurfaceArrays
from tests.helpers import GOLDEN, compat_params, read_T

pytestmark = pytest.mark.gpu
ERR_BUDGET = -5


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def both(lib, p, surf, moments=False, vrec=False):
    a = Buffers(p, 5, moments=moments, vrec=vrec)
    b = Buffers(p, 5, moments=moments, vrec=vrec)
    print(f"noalloc: mode {p.mode}, {p.nparticles} particles", flush=True)
    assert lib.dsmc_run_ballistic(C.byref(p), C.byref(surf.surface), C.byref(a.result)) == 0
    assert lib.dsmc_run_ballistic_noalloc(C.byref(p), C.byref(surf.surface), C.byref(b.result)) == 0
    return a, b


def same_particles(a, b):
    for k in ("steps", "end_state", "fate", "hops", "ttime"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert a.counters() == b.counters()


def test_noalloc_physical_equals_run_ballistic(lib):
    surf = W.icosphere(lib, 4, W.BODIES[0], 0.5)
    for n in (20000, 3000):
        p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=n, seed=8)
        p.subsolar_lon = 0.5
        a, b = both(lib, p, surf, moments=True)
        same_particles(a, b)
        assert np.array_equal(a.vcount, b.vcount) and np.array_equal(a.density, b.density)
        assert np.array_equal(a.horizon, b.horizon)
        cnt = np.repeat(a.vcount.astype(np.float64), 3)
        assert np.all(np.abs(a.vmom - b.vmom) <= 1e-12 * cnt * 5e3 + 1e-9)
        assert b.result.total_steps > 0 and b.result.elapsed_ms > 0


def test_noalloc_compat_equals_run_ballistic(lib):
    T, ftd, ffd = read_T(GOLDEN / "T_thermal.dat")
    surf = SurfaceArrays(T, ftd, ffd)
    for n in (1500, 300):
        p = compat_params(ftd, ffd, n, 4)
        a, b = both(lib, p, surf, vrec=True)
        same_particles(a, b)
        assert np.allclose(a.density, b.density, rtol=1e-12, atol=1e-12)
        assert np.allclose(a.horizon, b.horizon, rtol=1e-12, atol=1e-15)
        na, nb = int(a.result.vrec_n), int(b.result.vrec_n)
        assert na == nb > 0
        ca = np.bincount(a.vrec_td[:na] * p.hd + a.vrec_ht[:na], minlength=ftd * p.hd)
        cb = np.bincount(b.vrec_td[:nb] * p.hd + b.vrec_ht[:nb], minlength=ftd * p.hd)
        assert np.array_equal(ca, cb)


def test_noalloc_rejects_a_run_too_large(lib):
    """A run larger than the block is rejected; a compat run (transit times,
    records) before it and a physical run after it both equal
    dsmc_run_ballistic: nothing of one call's run reaches the next."""
    T, ftd, ffd = read_T(GOLDEN / "T_thermal.dat")
    pc = compat_params(ftd, ffd, 200, 6)
    a, b = both(lib, pc, SurfaceArrays(T, ftd, ffd), vrec=True)
    same_particles(a, b)
    surf = W.icosphere(lib, 3, W.BODIES[0], 0.0)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=2_000_000, seed=1)
    r = Buffers(p, 5, per_particle=False)
    print("noalloc: 2e6 particles (too large)", flush=True)
    assert lib.dsmc_run_ballistic_noalloc(C.byref(p), C.byref(surf.surface), C.byref(r.result)) == ERR_BUDGET
    p.nparticles = 500
    a, b = both(lib, p, surf)
    same_particles(a, b)
    assert not b.ttime.any()          # physical runs have no transit times
