"""Shared test helpers: oracle loading, reference presets, fixtures, .dat I/O."""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

from dsmc_suite_amd.capi import (MODE_COMPAT, Buffers, Params, Phase2Input, SurfaceArrays, ptr)

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
ORACLE_DIR = ROOT / "oracle"

ORNG_KEYED, ORNG_SEQUENTIAL, ORNG_RAND_R = 0, 1, 2


def oracle(devmath: bool = False) -> C.CDLL:
    """The oracle build: libm (bit-exact with the reference) or device-math."""
    name = "libcpu_ref_dev.so" if devmath else "libcpu_ref.so"
    path = ORACLE_DIR / name
    if not path.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR), "-s"], check=True)
    lib = C.CDLL(str(path))
    lib.cpu_ref_velocity_mean_randr.restype = C.c_double
    lib.cpu_ref_velocity_mean_randr.argtypes = [C.c_double, C.c_uint, C.c_int, C.c_double,
                                                C.c_double, C.c_double]
    lib.cpu_ref_sample_speed.restype = C.c_double
    lib.cpu_ref_sample_speed.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double,
                                         C.c_int, C.c_double, C.c_int]
    lib.cpu_ref_elementary.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.c_int64]
    return lib


def read_T(path) -> tuple[np.ndarray, int, int]:
    """Temperature .dat (fileio.c:47-62): ftd, ffd, then ftd*ffd values."""
    vals = Path(path).read_text().split()
    ftd, ffd = int(vals[0]), int(vals[1])
    return np.array(vals[2:2 + ftd * ffd], dtype=np.float64), ftd, ffd


def compat_params(ftd: int, ffd: int, n: int, seed: int, advanced: bool = False) -> Params:
    """The reference's hard-coded parameters: ballistic.c:49-67 / advanced.c:131-152,
    include/constants.h:5-35, caps at ballistic.c:77,216 and advanced.c:169,44."""
    p = Params()
    p.mode = MODE_COMPAT
    p.stream = 2 if advanced else 1
    p.seed = seed
    p.nparticles = n
    p.G, p.M, p.radius = 6.67e-11, 9.39e20, 476.2e3
    p.k_B, p.mmass, p.mole = 1.38e-23, 18e-3, 6.02e23
    p.lifetime = 1.0 / 1.02e-5 * (2.8 * 2.8)
    p.period = 9.07417 * 3600.0
    p.alpha = 0.5
    p.v_esc = math.sqrt(2.0 * p.G * p.M / p.radius)
    p.max_hop_steps, p.max_steps = 10000000, 0
    p.t_freeze, p.w_min = 100.0, 0.01
    p.ftd, p.ffd = ftd, ffd
    if not advanced:
        p.dt, p.hd, p.exp_grid = 5.0, 100, 0
        p.unih, p.alpha2 = p.radius * 0.1, 0.0
        p.max_height = p.unih * p.hd + p.radius
    else:
        p.dt, p.hd, p.exp_grid = 0.5, 50, 1
        p.max_height = 3e6 * math.sqrt(2.0)
        p.alpha2 = 0.04
        p.unih = (p.max_height - p.radius) / (math.pow(1.0 + p.alpha2, p.hd) - 1.0) * p.alpha2
    p.record_cap_cell = 200
    p.record_cap_total = 100 * ftd * p.hd
    p.cross_section = 1e-19
    p.fit_degree, p.max_per_cell, p.min_fit_samples = 3, 50, 8
    p.density_scale = 1.0
    p.sublimation_scale = 1.0
    return p


def golden(name: str) -> dict:
    z = np.load(GOLDEN / name, allow_pickle=False)
    return {k: z[k] for k in z.files}


def dense(idx, val, n) -> np.ndarray:
    a = np.zeros(n)
    a[idx] = val
    return a


def phase2(fx: dict) -> tuple[Phase2Input, list]:
    """Phase2Input from the advanced fixture's inputs; returns (struct, keepalive)."""
    den = np.ascontiguousarray(fx["in_den"], dtype=np.float64)
    vr = np.ascontiguousarray(fx["in_vrec"], dtype=np.float64).ravel()
    td = np.ascontiguousarray(fx["in_vrec_td"], dtype=np.int32)
    fd = np.ascontiguousarray(fx["in_vrec_fd"], dtype=np.int32)
    ht = np.ascontiguousarray(fx["in_vrec_ht"], dtype=np.int32)
    inp = Phase2Input()
    inp.copy_atm = ptr(den, C.c_double)
    inp.atm_ftd, inp.atm_ffd, inp.atm_hd = (int(x) for x in fx["in_den_dims"])
    inp.vrec = ptr(vr, C.c_double)
    inp.vrec_td, inp.vrec_fd, inp.vrec_ht = ptr(td, C.c_int32), ptr(fd, C.c_int32), ptr(ht, C.c_int32)
    inp.nrec = len(td)
    return inp, [den, vr, td, fd, ht]


def run_oracle(lib, p: Params, surf: SurfaceArrays, rows=5, nthreads=1, rng=ORNG_KEYED,
               vrec=False, moments=False, phase2_input=None) -> Buffers:
    b = Buffers(p, rows, vrec=vrec, moments=moments)
    if phase2_input is None:
        rc = lib.cpu_ref_ballistic(C.byref(p), C.byref(surf.surface), C.byref(b.result), nthreads, rng)
    else:
        rc = lib.cpu_ref_advanced(C.byref(p), C.byref(surf.surface), C.byref(phase2_input),
                                  C.byref(b.result), nthreads, rng)
    assert rc == 0, rc
    return b


def c_fmt(fmt: str, vals) -> str:
    """printf-style formatting of a sequence (Python's % matches C for these specs)."""
    return "".join(fmt % v for v in vals)


# ---------------------------------------------------------------- statistics
# Coarse cells of a [ht][td][fd] grid (36 x 72 angular cells): the surface shell
# in 12 latitude bands x 12 longitude sectors (15 x 30 degrees; most of the gas
# of the compat runs never leaves it), the shells above in groups that double in
# depth, 6 x 6 (30 x 60 degrees).
COARSE_HT = (0, 1, 2, 3, 4, 8, 16, 32, 100)


def coarse_density(den: np.ndarray, hd: int, ftd: int, ffd: int, ht_edges=COARSE_HT) -> np.ndarray:
    """Sums of a [ht][td][fd] density field over coarse cells, flattened."""
    d = np.asarray(den, dtype=np.float64).reshape(hd, ftd, ffd)
    surface = d[0].reshape(12, ftd // 12, 12, ffd // 12).sum(axis=(1, 3)).ravel()
    c = d.reshape(hd, 6, ftd // 6, 6, ffd // 6).sum(axis=(2, 4))
    edges = [e for e in ht_edges if 1 <= e <= hd]
    if edges[-1] != hd:
        edges.append(hd)
    above = [c[a:b].sum(axis=0).ravel() for a, b in zip(edges[:-1], edges[1:])]
    return np.concatenate([surface, *above])


def two_sample_chi2(a: np.ndarray, b: np.ndarray, edges=None) -> tuple[float, int, float]:
    """Two-sample chi-squared test of two samples of one quantity: X^2 =
    sum (sqrt(Nb/Na) a_k - sqrt(Na/Nb) b_k)^2 / (a_k + b_k) over bins, with
    (bins - 1) degrees of freedom.  Bins: `edges` (right-open), or the pooled
    quantiles with ~25 values of the smaller sample per bin (4 to 16 bins); bins
    with fewer than 10 pooled values are merged into their neighbour.  Returns
    (X^2, dof, p)."""
    from scipy.stats import chi2
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    pooled = np.concatenate([a, b])
    if edges is None:
        nbins = int(np.clip(min(len(a), len(b)) // 25, 4, 16))
        edges = np.unique(np.quantile(pooled, np.linspace(0, 1, nbins + 1)))
        edges[-1] = np.nextafter(edges[-1], np.inf)      # the largest value in the last bin
    edges = np.asarray(edges, np.float64)
    ca = np.histogram(a, edges)[0].astype(np.float64)
    cb = np.histogram(b, edges)[0].astype(np.float64)
    # merge sparse bins
    ma, mb, acc_a, acc_b = [], [], 0.0, 0.0
    for x, y in zip(ca, cb):
        acc_a += x
        acc_b += y
        if acc_a + acc_b >= 10:
            ma.append(acc_a); mb.append(acc_b)
            acc_a = acc_b = 0.0
    if acc_a + acc_b > 0:
        if ma:
            ma[-1] += acc_a; mb[-1] += acc_b
        else:
            ma.append(acc_a); mb.append(acc_b)
    ca, cb = np.array(ma), np.array(mb)
    na, nb = ca.sum(), cb.sum()
    x2 = float(np.sum((np.sqrt(nb / na) * ca - np.sqrt(na / nb) * cb) ** 2 / (ca + cb)))
    dof = len(ca) - 1
    return x2, dof, float(chi2.sf(x2, dof)) if dof > 0 else 1.0


def per_cell_t(ref: np.ndarray, reps: np.ndarray, frac_of_max: float = 1e-3) -> tuple[float, int]:
    """One reference sample per coarse cell against R replicas of the same size:
    t_c = (ref_c - mean_c) / (s_c sqrt(1 + 1/R)) ~ Student t(R - 1) per cell;
    returns the smallest two-sided p over the cells whose replica mean exceeds
    frac_of_max of the largest one, Bonferroni-adjusted (times the number of
    cells: valid whatever the correlation between cells), and that number."""
    from scipy.stats import t as student
    reps = np.asarray(reps, np.float64)
    R = reps.shape[0]
    m, s = reps.mean(axis=0), reps.std(axis=0, ddof=1)
    keep = (m > frac_of_max * m.max()) & (s > 0)
    tc = (np.asarray(ref)[keep] - m[keep]) / (s[keep] * math.sqrt(1.0 + 1.0 / R))
    p = 2.0 * student.sf(np.abs(tc), R - 1)
    return float(min(1.0, p.min() * keep.sum())), int(keep.sum())
