"""Physical collisional runs: phase-1 speed histograms, partner-speed fits,
number density, phase-2 collisions.

The reference has no physical collisional mode (its advanced.c is the compat
mode, tested in test_oracle_reference.py / test_gpu_parity.py); this mode's
oracle functions are pinned here by what the mathematics fixes:
  fits            numpy's least-squares polyfit on the same points (QR against
                  the oracle's normal equations: rtol 1e-6 on the fitted values);
  number density  the closed-form shell-sector volume and emission rate;
  scattering      invariants of an elastic equal-mass collision (relative speed
                  and centre-of-mass velocity conserved, isotropic outcome);
  histograms      brute force: the visit rule restated in numpy on a
                  single-particle trajectory is not available, so they are
                  pinned by the count identity (sum over bins = deposit count per
                  cell) and by the lane state machine through every record path.
The GPU tests compare the HIP path with the oracle (device-math build) bit for
bit per particle, histograms and integer counts exactly, moments to 1e-12.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from dsmc_suite_amd import capi
from dsmc_suite_amd import workloads as W
from dsmc_suite_amd.capi import PRESET_CERES_PHYSICAL_COLL, Buffers, Phase2Input, ptr
from tests.helpers import oracle

LEVELS = 3


@pytest.fixture(scope="module")
def lib():
    return capi.load()


@pytest.fixture(scope="module")
def emu():
    import pathlib
    lib = C.CDLL(str(pathlib.Path(__file__).parent / "emu" / "liblane_emu.so"))
    lib.emu_ballistic_path.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int64]
    lib.emu_physical_coll.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int, C.c_int64]
    return lib


def orc(devmath=True):
    o = oracle(devmath=devmath)
    o.cpu_ref_sample_speed_phys.restype = C.c_double
    o.cpu_ref_sample_speed_phys.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double, C.c_double]
    return o


def coll_params(lib, n, seed, scale=1e4, stream=1, hd=30):
    """Ceres collisional preset on a smaller grid; sublimation scaled up so that
    a few thousand test particles see collisions."""
    p = capi.preset(lib, PRESET_CERES_PHYSICAL_COLL, 36, 72, n=n, seed=seed)
    p.hd = hd
    p.unih = (p.max_height - p.radius) / ((1.0 + p.alpha2) ** hd - 1.0) * p.alpha2
    p.sublimation_scale = scale
    p.stream = stream
    p.vhist_bins = 48
    p.vhist_max = 1500.0
    return p


def surface(lib):
    return W.icosphere(lib, LEVELS, W.BODIES[0], 0.0)


def phase1(lib, o, p, surf, nthreads=4):
    b = Buffers(p, 5, per_particle=True, moments=True, vhist=True)
    assert o.cpu_ref_ballistic(C.byref(p), C.byref(surf.surface), C.byref(b.result), nthreads, 0) == 0
    return b


def fits(o, p, vhist):
    cells = p.ftd * p.ffd * p.hd
    cf = np.zeros(cells * (p.fit_degree + 1))
    lo, hi = np.zeros(cells), np.zeros(cells)
    valid = np.zeros(cells, np.int32)
    o.cpu_ref_hist_fits(C.byref(p), ptr(vhist, C.c_int64), ptr(cf, C.c_double), ptr(lo, C.c_double),
                        ptr(hi, C.c_double), ptr(valid, C.c_int32))
    return cf, lo, hi, valid


def density(o, p, surf, vcount):
    nd = np.zeros(p.ftd * p.ffd * p.hd)
    assert o.cpu_ref_number_density(C.byref(p), C.byref(surf.surface), ptr(vcount, C.c_int64),
                                    ptr(nd, C.c_double)) == 0
    return nd


# ---------------------------------------------------------------- oracle pins (CPU)
def test_hist_fits_match_numpy_polyfit(lib):
    o = orc(False)
    p = coll_params(lib, 10, 1, hd=2)
    p.ftd, p.ffd = 2, 3
    bins, deg = p.vhist_bins, p.fit_degree
    rng = np.random.default_rng(3)
    cells = p.ftd * p.ffd * p.hd
    vh = np.zeros((cells, bins), np.int64)
    for c in range(cells):
        lo = int(rng.integers(0, 10))
        hi = lo + int(rng.integers(8, 30))
        vh[c, lo:hi + 1] = rng.integers(1, 5000, hi - lo + 1)
    vh[0] = 0                                   # empty cell
    vh[1] = 0; vh[1, 5:9] = 3                  # too few bins for degree 6
    cf, pmin, pmax, valid = fits(o, p, vh.ravel())
    dv = p.vhist_max / bins
    assert valid[0] == 0 and valid[1] == 0
    for c in range(2, cells):
        h = vh[c]
        nz = np.flatnonzero(h)
        lo, hi = nz[0], nz[-1]
        npts = hi - lo + 1
        x = (np.arange(npts) + 0.5) / npts
        y = 50.0 * np.cumsum(h[lo:hi + 1][::-1])[::-1] / h.sum()
        want = np.polyfit(x, y, deg)
        got = cf[c * (deg + 1):(c + 1) * (deg + 1)]
        assert valid[c] == 1
        assert np.allclose(np.polyval(got, x), np.polyval(want, x), rtol=1e-6, atol=1e-6)
        assert pmin[c] == lo * dv and pmax[c] == (hi + 1) * dv


def test_sample_speed_phys_inverts_the_fit(lib):
    """The sampled speed solves P(x) = ranking to the 0.01 m/s bracket."""
    o = orc(False)
    cf = np.array([0.0, 0.0, 0.0, 0.0, 0.0, -50.0, 50.0])     # P(x) = 50 (1 - x)
    lo, hi = 100.0, 900.0
    for u in (0.1, 0.37, 0.5, 0.93):
        v = o.cpu_ref_sample_speed_phys(ptr(cf, C.c_double), 6, lo, hi, u)
        x = (v - lo) / (hi - lo)
        assert abs(50 * (1 - x) - u * 50) <= 50 * 0.01 / (hi - lo) + 1e-12


def test_number_density_closed_form(lib):
    o = orc(False)
    surf = surface(lib)
    p = coll_params(lib, 1000, 1, hd=20)
    p.production_rate = 0.0                     # the emission rate of the face weights
    rng = np.random.default_rng(5)
    cells = p.ftd * p.ffd * p.hd
    vcount = rng.integers(0, 10_000, cells).astype(np.int64)
    nd = density(o, p, surf, vcount)
    R = p.radius
    F = surf.face_w.sum() * R * R * p.sublimation_scale
    k = np.arange(p.hd + 1)
    r = R + p.unih * ((1 + p.alpha2) ** k - 1) / p.alpha2
    th = -np.pi / 2 + np.pi * np.arange(p.ftd + 1) / p.ftd
    V = ((r[1:] ** 3 - r[:-1] ** 3) / 3)[:, None, None] * np.diff(np.sin(th))[None, :, None] \
        * (2 * np.pi / p.ffd) * np.ones((1, 1, p.ffd))
    want = vcount.reshape(p.hd, p.ftd, p.ffd) * p.dt * F / (p.nparticles * V)
    assert np.allclose(nd, want.ravel(), rtol=1e-12, atol=0)
    q = p.copy()
    q.production_rate = 3.5e26                  # an explicit emission rate replaces the face sum
    nd2 = density(o, q, surf, vcount)
    assert np.allclose(nd2, nd * (3.5e26 * q.sublimation_scale) / F, rtol=1e-12, atol=0)
    got = np.zeros(cells)
    assert lib.dsmc_number_density(C.byref(p), C.byref(surf.surface), ptr(vcount, C.c_int64),
                                   ptr(got, C.c_double)) == 0
    assert np.array_equal(got, nd)              # product host code == oracle, bit for bit


def test_collision_kinematics_invariants(lib):
    o = orc(False)
    o.cpu_ref_collide_kinematics.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rng = np.random.default_rng(9)
    dirs = []
    for _ in range(4000):
        v = rng.normal(0, 400, 3)
        vv2 = float(rng.uniform(10, 900))
        u = rng.uniform(0, 1, 4)
        out, v2 = np.zeros(3), np.zeros(3)
        o.cpu_ref_collide_kinematics(ptr(v, C.c_double), vv2, ptr(u, C.c_double), ptr(out, C.c_double),
                                     ptr(v2, C.c_double))
        assert abs(np.linalg.norm(v2) - vv2) <= 1e-12 * vv2
        cm = 0.5 * (v + v2)
        g, g2 = np.linalg.norm(v - v2), 2 * np.linalg.norm(out - cm)
        assert abs(g - g2) <= 1e-9 * max(g, 1.0)           # relative speed conserved
        dirs.append((out - cm) / np.linalg.norm(out - cm))
    d = np.array(dirs)
    assert np.all(np.abs(d.mean(axis=0)) < 0.05)           # isotropic scattering
    assert abs((d[:, 2] ** 2).mean() - 1 / 3) < 0.03


def test_histogram_count_identity(lib):
    """Every deposit step lands in exactly one speed bin of its cell."""
    o = orc(True)
    surf = surface(lib)
    p = coll_params(lib, 400, 21)
    b = phase1(lib, o, p, surf)
    vh = b.vhist.reshape(-1, p.vhist_bins)
    assert np.array_equal(vh.sum(axis=1), b.vcount)
    assert b.vcount.sum() > 0


# ---------------------------------------------------------------- lane state machine vs oracle (CPU)
@pytest.mark.parametrize("path,budget", [(0, 0), (1, 23), (2, 23), (2, 4096), (3, 23), (4, 23)])
def test_phase1_histograms_through_record_paths(lib, emu, path, budget):
    o = orc(True)
    surf = surface(lib)
    p = coll_params(lib, 150, 31)
    want = phase1(lib, o, p, surf)
    got = Buffers(p, 5, per_particle=True, moments=True, vhist=True)
    assert emu.emu_ballistic_path(C.byref(p), C.byref(surf.surface), C.byref(got.result), path, budget) == 0
    for k in ("vhist", "vcount", "steps", "end_state", "fate"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k


@pytest.mark.parametrize("path,budget", [(0, 0), (2, 17), (2, 1024)])
def test_phase2_lane_state_machine_matches_oracle(lib, emu, path, budget):
    o = orc(True)
    surf = surface(lib)
    p1 = coll_params(lib, 1500, 41)
    b1 = phase1(lib, o, p1, surf, nthreads=8)
    nd = density(o, p1, surf, b1.vcount)
    p2 = coll_params(lib, 120, 41, stream=2)
    cf, pmin, pmax, valid = fits(o, p2, b1.vhist)
    nsig = np.where(valid == 1, nd * p2.cross_section, 0.0)
    inp = Phase2Input()
    inp.ndens = ptr(nd, C.c_double)
    inp.vhist = ptr(b1.vhist, C.c_int64)
    want = Buffers(p2, 5, per_particle=True, moments=True)
    assert o.cpu_ref_advanced(C.byref(p2), C.byref(surf.surface), C.byref(inp), C.byref(want.result), 4, 0) == 0
    got = Buffers(p2, 5, per_particle=True, moments=True)
    assert emu.emu_physical_coll(C.byref(p2), C.byref(surf.surface), nsig.ctypes.data, cf.ctypes.data,
                                 pmin.ctypes.data, pmax.ctypes.data, C.byref(got.result), path, budget) == 0
    for k in ("steps", "end_state", "fate", "hops", "vcount"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    assert want.result.n_collisions > 20, want.result.n_collisions


# ---------------------------------------------------------------- HIP path vs oracle (GPU)
@pytest.mark.gpu
def test_gpu_phase1_histograms_and_fits(lib):
    o = orc(True)
    surf = surface(lib)
    p = coll_params(lib, 3000, 51)
    want = phase1(lib, o, p, surf, nthreads=8)
    got = Buffers(p, 5, per_particle=True, moments=True, vhist=True)
    assert lib.dsmc_run_ballistic(C.byref(p), C.byref(surf.surface), C.byref(got.result)) == 0
    for k in ("vhist", "vcount", "steps", "end_state", "fate", "hops"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    cf, pmin, pmax, valid = fits(o, p, want.vhist)
    cells = p.ftd * p.ffd * p.hd
    gcf, glo, ghi = np.zeros_like(cf), np.zeros(cells), np.zeros(cells)
    gvalid = np.zeros(cells, np.int32)
    assert lib.dsmc_build_hist_fits(C.byref(p), ptr(want.vhist, C.c_int64), ptr(gcf, C.c_double),
                                    ptr(glo, C.c_double), ptr(ghi, C.c_double), ptr(gvalid, C.c_int32)) == 0
    assert np.array_equal(gvalid, valid) and valid.sum() > 50
    assert np.array_equal(gcf, cf) and np.array_equal(glo, pmin) and np.array_equal(ghi, pmax)


@pytest.mark.gpu
def test_gpu_collisional_pipeline_matches_oracle(lib):
    o = orc(True)
    o.cpu_ref_collisional.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_int]
    surf = surface(lib)
    p1 = coll_params(lib, 2000, 61)
    p2 = coll_params(lib, 800, 61, stream=2)
    w1 = Buffers(p1, 5, per_particle=False, moments=True, vhist=True)
    w2 = Buffers(p2, 5, per_particle=True, moments=True)
    assert o.cpu_ref_collisional(C.byref(p1), C.byref(p2), C.byref(surf.surface), C.byref(w1.result),
                                 C.byref(w2.result), 8, 0) == 0
    g1 = Buffers(p1, 5, per_particle=False, moments=True, vhist=True)
    g2 = Buffers(p2, 5, per_particle=True, moments=True)
    print(f"collisional pipeline p1 n={p1.nparticles} p2 n={p2.nparticles}", flush=True)
    assert lib.dsmc_run_collisional(C.byref(p1), C.byref(p2), C.byref(surf.surface), C.byref(g1.result),
                                    C.byref(g2.result)) == 0
    assert np.array_equal(g1.vhist, w1.vhist) and np.array_equal(g1.vcount, w1.vcount)
    for k in ("steps", "end_state", "fate", "hops", "vcount"):
        assert np.array_equal(getattr(g2, k), getattr(w2, k)), k
    assert g2.counters() == w2.counters()
    assert w2.result.n_collisions > 50
    cnt = np.repeat(w2.vcount.astype(np.float64), 3)
    assert np.all(np.abs(g2.vmom - w2.vmom) <= 1e-12 * cnt * 5e3 + 1e-9)
