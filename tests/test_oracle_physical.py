"""Independent pins of the oracle's physical mode (oracle/cpu_ref.c).

The benchmark configurations run the physical exosphere model, which the
reference does not contain; the HIP kernels equal the oracle's device-math build
bit for bit (tests/test_gpu_parity.py), so what pins the model is checked here,
on the oracle itself, against facts that do not come from its code:

* launch weights = flat face area x Hertz-Knudsen flux P(T) / sqrt(2 pi m k T)
  with the Antoine coefficients of the reference's thermal model
  (src/thermal/thermal_model.c:33-38,50-55), and that law against an
  independent ice vapour-pressure fit (Murphy & Koop 2005);
* the launch-face histogram against those weights (chi-squared);
* launch points uniform on the flat face (barycentric coordinates Dirichlet(1,1,1));
* emission: flux-weighted Maxwellian speed (m v^2 / 2kT ~ Gamma(2)), cosine-law
  (Lambertian) direction (cos^2 theta ~ U(0,1)), uniform azimuth;
* Monte-Carlo photolysis: the particle ends on the sunlit step where its sunlit
  count reaches ceil(E tau / dt), E ~ Exp(1) (geometric law of the step count);
* face location: the face returned holds the point (all three edge margins >= 0)
  and is the face an exhaustive search over every finest face picks.

Each check is paired with a negative control: the same statistic applied to what
a plausible mistake (a dropped square root or temperature factor, a changed
coefficient, a factor 2 in the variance, an off-by-one step) would produce, which
must fail.  Both oracle builds (libm and device-math) are checked.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

from dsmc_suite_amd import capi
from dsmc_suite_amd import workloads as W
from dsmc_suite_amd.capi import PRESET_CERES_PHYSICAL, Buffers, SurfaceArrays, ptr
from tests.helpers import ROOT, oracle, run_oracle

K_B = W.K_B
REF_THERMAL = ROOT / "oracle" / "_ref" / "libref_thermal.so"
DP = C.POINTER(C.c_double)


def bind(o):
    o.cpu_ref_phys_launch.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int32), DP, DP,
                                      C.POINTER(C.c_int64)]
    o.cpu_ref_locate_face.argtypes = [C.c_void_p, C.c_void_p, DP, C.c_int64, C.POINTER(C.c_int32)]
    return o


@pytest.fixture(scope="module")
def lib():
    return capi.load()


@pytest.fixture(scope="module", params=[False, True], ids=["libm", "devmath"])
def orc(request):
    return bind(oracle(devmath=request.param))


def antoine_coefficients(lib):
    """p_mat and the Torr->Pa factor of the thermal model's defaults: this
    library's (pinned against the reference's thermal_solve bit for bit) and,
    where it is built here, the compiled reference's own."""
    tp = capi.ThermalParams()
    lib.thermal_params_default(C.byref(tp))
    coef = (list(tp.p_mat), tp.tran_unit_p)
    if REF_THERMAL.exists():
        ref = C.CDLL(str(REF_THERMAL))
        rp = capi.ThermalParams()
        ref.thermal_params_default(C.byref(rp))
        assert (list(rp.p_mat), rp.tran_unit_p) == coef
    return coef


def antoine(T, coef):
    p, unit = coef
    T = np.asarray(T, float)
    return unit * 10.0 ** (p[0] / T + p[1] + p[2] * T + p[3] * T * T + p[4] * np.log10(T))


def flat_areas(surf: SurfaceArrays):
    nf = len(surf.face_T)
    f = surf.faces.reshape(-1, 3)[-nf:]
    v = surf.vert.reshape(-1, 3)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1), (a, b, c)


def launches(o, p, surf, n):
    face = np.zeros(n, np.int32)
    pos, vel, nph = np.zeros(n * 3), np.zeros(n * 3), np.zeros(n, np.int64)
    assert o.cpu_ref_phys_launch(C.byref(p), C.byref(surf.surface), n, ptr(face, C.c_int32), ptr(pos, C.c_double),
                                 ptr(vel, C.c_double), ptr(nph, C.c_int64)) == 0
    return face, pos.reshape(n, 3), vel.reshape(n, 3), nph


def grouped_chi2(counts, expected, min_expected=20.0):
    """Chi-squared p-value with cells merged (in order of expectation) until each
    holds at least `min_expected`."""
    order = np.argsort(expected)
    oc, ec, o_acc, e_acc = [], [], 0.0, 0.0
    for i in order:
        o_acc += counts[i]
        e_acc += expected[i]
        if e_acc >= min_expected:
            oc.append(o_acc); ec.append(e_acc)
            o_acc = e_acc = 0.0
    if e_acc > 0:
        oc[-1] += o_acc; ec[-1] += e_acc
    oc, ec = np.array(oc), np.array(ec)
    return stats.chisquare(oc, ec * oc.sum() / ec.sum()).pvalue


# ---------------------------------------------------------------- launch weights
@pytest.mark.parametrize("levels,body,lon", [(3, 0, 0.0), (4, 0, 1.1), (4, 3, 4.0), (2, 2, 2.5)])
def test_face_weights_are_antoine_flux_times_area(lib, levels, body, lon):
    b = W.BODIES[body]
    surf = W.icosphere(lib, levels, b, lon)
    area, (a, bb, c) = flat_areas(surf)
    # temperature at the normalised centroid, sun in the equator at `lon`
    cen = a + bb + c
    mu = (cen[:, 0] * math.sin(lon) + cen[:, 1] * math.cos(lon)) / np.linalg.norm(cen, axis=1)
    T = b.t_night + (b.t_day - b.t_night) * np.maximum(mu, 0.0) ** 0.25
    assert np.allclose(surf.face_T, T, rtol=1e-14, atol=0)
    coef = antoine_coefficients(lib)
    want = area * antoine(T, coef) / np.sqrt(2 * math.pi * b.mass * K_B * T)
    assert np.allclose(surf.face_w, want, rtol=1e-12, atol=0)
    # negative controls: a dropped sqrt(T) in the flux, a changed coefficient
    assert not np.allclose(surf.face_w, area * antoine(T, coef) / np.sqrt(2 * math.pi * b.mass * K_B), rtol=1e-3)
    bad = ([coef[0][0] * 1.001] + coef[0][1:], coef[1])
    assert not np.allclose(surf.face_w, area * antoine(T, bad) / np.sqrt(2 * math.pi * b.mass * K_B * T), rtol=1e-3)


def test_antoine_law_against_an_independent_ice_fit(lib):
    """The Antoine law the weights use (thermal_model.c:33-38,50-55) against the
    vapour pressure of hexagonal ice of Murphy & Koop (2005, Q. J. R. Meteorol.
    Soc. 131, eq. 7), an independent fit: within 1% over 200-260 K, within 25%
    down to 130 K (the Antoine fit drifts low in the cold end: 3.5% at 180 K,
    21% at 130 K)."""
    coef = antoine_coefficients(lib)
    T = np.linspace(130.0, 260.0, 27)
    mk = np.exp(9.550426 - 5723.265 / T + 3.53068 * np.log(T) - 0.00728332 * T)
    dev = np.abs(antoine(T, coef) / mk - 1.0)
    assert np.all(dev[T >= 200.0] < 0.01) and np.all(dev < 0.25)
    assert 0.14 < antoine(200.0, coef) < 0.19                 # Pa, ice at 200 K
    # a wrong sign of the 1/T term is off by orders of magnitude
    bad = ([-coef[0][0]] + coef[0][1:], coef[1])
    assert np.all(np.abs(antoine(T, bad) / mk - 1.0) > 1.0)


def test_launch_faces_follow_the_weights(lib, orc):
    """Launch-face histogram of 1.5M oracle launches against the analytic
    weights (chi-squared, cells of >= 20 expected); launches drawn with weights
    of a plausible mistake (a dropped sqrt(T) in the flux, a natural log for the
    log10 of the Antoine law) fail the same test against the right weights."""
    b = W.BODIES[0]
    surf = W.icosphere(lib, 3, b, 0.7)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=0, seed=11)
    p.subsolar_lon = 0.7
    n = 1500000
    face, *_ = launches(orc, p, surf, n)
    nf = len(surf.face_w)
    counts = np.bincount(face, minlength=nf)
    assert counts.sum() == n and face.min() >= 0 and face.max() < nf
    assert counts[surf.face_w == 0].sum() == 0
    assert grouped_chi2(counts, surf.face_w / surf.face_w.sum() * n) > 0.01
    area, _ = flat_areas(surf)
    coef = antoine_coefficients(lib)
    T = surf.face_T
    pm, unit = coef
    ln_law = unit * 10.0 ** (pm[0] / T + pm[1] + pm[2] * T + pm[3] * T * T + pm[4] * np.log(T))
    for wrong in (area * antoine(T, coef) / np.sqrt(2 * math.pi * b.mass * K_B),
                  area * ln_law / np.sqrt(T)):
        bad = SurfaceArrays(levels=3, vert=surf.vert, faces=surf.faces, face_T=surf.face_T, face_w=wrong)
        fb, *_ = launches(orc, p, bad, n)
        cb = np.bincount(fb, minlength=nf)
        assert grouped_chi2(cb, surf.face_w / surf.face_w.sum() * n) < 1e-6


# ---------------------------------------------------------------- launch point
def test_launch_points_uniform_on_the_face(lib, orc):
    """The launch point, taken back from radius R along its direction to the flat
    face, has barycentric coordinates inside the face and Dirichlet(1,1,1)
    distributed (uniform on the triangle): each coordinate ~ Beta(1, 2) and
    l_b / (1 - l_a) ~ U(0, 1).  A dropped square root of the barycentric map
    (r1 = u1) fails."""
    b = W.BODIES[0]
    surf = W.icosphere(lib, 2, b, 0.0)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=0, seed=5)
    n = 200000
    face, pos, _, _ = launches(orc, p, surf, n)
    assert np.allclose(np.linalg.norm(pos, axis=1), p.radius, rtol=1e-14)
    _, (a, bb, c) = flat_areas(surf)
    a, bb, c = a[face], bb[face], c[face]
    nrm = np.cross(bb - a, c - a)
    pt = pos * (np.sum(nrm * a, axis=1) / np.sum(nrm * pos, axis=1))[:, None]
    # barycentric coordinates by areas
    tot = np.linalg.norm(nrm, axis=1)
    la = np.sum(np.cross(bb - pt, c - pt) * nrm, axis=1) / tot ** 2
    lb = np.sum(np.cross(c - pt, a - pt) * nrm, axis=1) / tot ** 2
    lc = 1.0 - la - lb
    assert min(la.min(), lb.min(), lc.min()) > -1e-12
    beta = stats.beta(1, 2).cdf
    for lam in (la, lb, lc):
        assert stats.kstest(lam, beta).pvalue > 0.01
    assert stats.kstest(lb / (1.0 - la), "uniform").pvalue > 0.01
    # the same draws through r1 = u1 instead of sqrt(u1): l_a' = 1 - (1 - l_a)^2
    assert stats.kstest(1.0 - (1.0 - la) ** 2, beta).pvalue < 1e-6


# ---------------------------------------------------------------- emission
@pytest.mark.parametrize("body", [0, 3])
def test_emission_is_flux_maxwellian_and_lambertian(lib, orc, body):
    b = W.BODIES[body]
    surf = W.icosphere(lib, 3, b, 2.0)
    p = W.body_params(lib, b, 0, 2.0, 9)
    n = 200000
    face, pos, vel, _ = launches(orc, p, surf, n)
    omega = 2 * math.pi / p.period
    rel = vel - np.stack([omega * pos[:, 1], -omega * pos[:, 0], np.zeros(n)], axis=1)
    nrm = pos / np.linalg.norm(pos, axis=1)[:, None]
    speed = np.linalg.norm(rel, axis=1)
    T = surf.face_T[face]
    w = b.mass * speed ** 2 / (2 * K_B * T)
    cos_t = np.sum(rel * nrm, axis=1) / speed
    assert cos_t.min() > 0.0
    gamma2 = stats.gamma(2).cdf
    assert stats.kstest(w, gamma2).pvalue > 0.01                       # flux-weighted speed
    assert stats.kstest(cos_t ** 2, "uniform").pvalue > 0.01            # cosine law
    hxy = np.hypot(nrm[:, 0], nrm[:, 1])
    e1 = np.stack([-nrm[:, 1] / hxy, nrm[:, 0] / hxy, np.zeros(n)], axis=1)
    e2 = np.cross(nrm, e1)
    phi = np.arctan2(np.sum(rel * e2, axis=1), np.sum(rel * e1, axis=1)) % (2 * math.pi)
    assert stats.kstest(phi / (2 * math.pi), "uniform").pvalue > 0.01   # azimuth
    # negative controls: a volume (not flux) Maxwellian would have w ~ Gamma(1.5);
    # sigma^2 = 2kT/m doubles w; a dropped sqrt in cos_t = sqrt(u) gives cos_t^2
    assert stats.kstest(w, stats.gamma(1.5).cdf).pvalue < 1e-6
    assert stats.kstest(2 * w, gamma2).pvalue < 1e-6
    assert stats.kstest(cos_t ** 4, "uniform").pvalue < 1e-6


# ---------------------------------------------------------------- photolysis
def test_photolysis_step_law(lib, orc):
    """Without gravity, particles launched from the subsolar face fly straight
    out; those moving sunward stay sunlit on every step, so each must end
    photolysed on exactly its ceil(E tau / dt)-th step, and those step counts
    follow the geometric law P(steps >= k) = exp(-(k - 1) dt / tau).  An
    off-by-one law (floor) or a wrong tau / dt fails."""
    b = W.BODIES[0]
    lon = 0.9
    surf = W.icosphere(lib, 3, b, lon)
    hot = int(np.argmax(surf.face_T))
    w = np.zeros_like(surf.face_w)
    w[hot] = 1.0
    one = SurfaceArrays(levels=3, vert=surf.vert, faces=surf.faces, face_T=surf.face_T, face_w=w)
    n = 20000
    p = W.body_params(lib, b, n, lon, 21)
    p.M = 0.0                      # G * M = 0: straight lines
    p.period = 1e30                # no rotation velocity
    p.lifetime = 2.0e4             # tau / dt = 2000 steps
    p.max_height = 1e4 * p.radius
    p.max_steps = 0
    p.unih = (p.max_height - p.radius) / ((1.0 + p.alpha2) ** p.hd - 1.0) * p.alpha2
    res = run_oracle(orc, p, one, nthreads=8)
    face, pos, vel, nph = launches(orc, p, one, n)
    assert np.all(face == hot)
    sun = np.array([math.sin(lon), math.cos(lon)])
    fin_v = res.end_state.reshape(n, 6)[:, 3:]
    assert np.array_equal(fin_v, vel)                        # no force: the launch velocity
    sunward = (vel[:, :2] @ sun >= 0) & (pos[:, :2] @ sun > 0)
    assert sunward.sum() > n // 3
    assert np.all(res.fate[sunward] == 3)
    assert np.array_equal(res.steps[sunward], nph[sunward])
    assert not np.array_equal(res.steps[sunward], nph[sunward] - 1)
    a = p.dt / p.lifetime
    edges = np.unique(np.round(np.quantile(nph, np.linspace(0, 1, 41))).astype(np.int64))
    edges[-1] = nph.max() + 1
    obs = np.histogram(nph, bins=edges)[0]
    surv = lambda k: np.exp(-(np.maximum(k, 1) - 1) * a)        # P(n >= k)
    exp_p = surv(edges[:-1]) - surv(edges[1:])
    assert stats.chisquare(obs, exp_p / exp_p.sum() * n).pvalue > 0.01
    for f in (1.1, 2.0):             # a rate 10% off; dt / tau with the collisional dt of 5 s
        wrong = lambda k: np.exp(-(np.maximum(k, 1) - 1) * a * f)
        exp_w = wrong(edges[:-1]) - wrong(edges[1:])
        assert stats.chisquare(obs, exp_w / exp_w.sum() * n).pvalue < 1e-6


# ---------------------------------------------------------------- face location
@pytest.mark.parametrize("levels", [0, 1, 2])
def test_locate_face_holds_the_point(lib, orc, levels):
    surf = W.icosphere(lib, levels, W.BODIES[0], 0.0)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=0, seed=0)
    nf = len(surf.face_T)
    ff = surf.faces.reshape(-1, 3)[-nf:]
    v = surf.vert.reshape(-1, 3)
    rng = np.random.default_rng(levels)
    d = [rng.normal(size=(20000, 3)), v]
    ea, eb = v[ff[:, 0]], v[ff[:, 1]]
    mid = 0.5 * (ea + eb)
    d.append(mid)
    perp = np.cross(eb - ea, mid)
    perp /= np.linalg.norm(perp, axis=1)[:, None]
    for off in (1e-13, 1e-11, 1e-9):
        d += [mid + off * perp, mid - off * perp]
    d = np.ascontiguousarray(np.concatenate(d))
    out = np.zeros(len(d), np.int32)
    assert orc.cpu_ref_locate_face(C.byref(p), C.byref(surf.surface), ptr(d, C.c_double), len(d),
                                   ptr(out, C.c_int32)) == 0
    a, b, c = v[ff[:, 0]], v[ff[:, 1]], v[ff[:, 2]]
    nab, nbc, nca = np.cross(a, b), np.cross(b, c), np.cross(c, a)

    def margins(dd, f):
        return np.stack([np.sum(dd * nab[f], 1), np.sum(dd * nbc[f], 1), np.sum(dd * nca[f], 1)], 1).min(1)

    tol = 1e-14 * np.linalg.norm(d, axis=1)
    assert np.all(margins(d, out) >= -tol)
    # exhaustive search: the face with the largest minimum margin
    full = np.stack([np.minimum(np.minimum(d @ nab.T[:, f:f + 1], d @ nbc.T[:, f:f + 1]),
                                d @ nca.T[:, f:f + 1])[:, 0] for f in range(nf)], 1)
    best = full.argmax(1)
    same = best == out
    assert np.all(same | (margins(d, best) <= tol))          # differ only on an edge or vertex
    assert same[:20000].mean() > 0.999
    # negative control: the neighbouring index (a plausible off-by-one) does not hold the points
    assert np.mean(margins(d[:20000], (out[:20000] + 1) % nf) >= -tol[:20000]) < 0.05
