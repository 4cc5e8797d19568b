"""Synthetic multi-run workloads: bodies, surfaces and parameter sets.

Four exosphere bodies (GM, radius, species mass, synthetic surface temperature
range T = t_night + (t_day - t_night) * max(cos sza, 0)^0.25 on an icosphere),
and generators for the batched workload (many independent ballistic runs with
log-uniform particle counts, random body and subsolar longitude) and for a
Poisson request stream of the same kind.  Everything is seeded; nothing here
computes a run - the runs go through the C ABI (capi).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from .capi import PRESET_CERES_PHYSICAL, Params, SurfaceArrays, preset, ptr

K_B = 1.380649e-23
CERES_PRODUCTION = 2.0e26     # molecules/s, the order of Ceres' measured water outgassing


@dataclass(frozen=True)
class Body:
    name: str
    GM: float        # m^3 s^-2
    radius: float    # m
    mass: float      # kg per molecule
    t_night: float   # K
    t_day: float     # K
    period: float    # s, rotation


BODIES = (
    Body("Ceres H2O", 6.26e10, 470e3, 2.99e-26, 100.0, 235.0, 9.07417 * 3600.0),
    Body("Moon Na", 4.90e12, 1737e3, 3.82e-26, 100.0, 390.0, 27.321661 * 86400.0),
    Body("Europa H2O", 3.20e12, 1561e3, 2.99e-26, 80.0, 130.0, 3.551181 * 86400.0),
    Body("Mercury Na", 2.20e13, 2440e3, 3.82e-26, 100.0, 700.0, 58.646 * 86400.0),
)


def icosphere(lib, levels: int, body: Body, lon: float) -> SurfaceArrays:
    """Icosphere surface of `body` with the sun at longitude `lon` [rad]."""
    nv, na, nf = C.c_int32(), C.c_int32(), C.c_int32()
    lib.dsmc_icosphere_sizes(levels, C.byref(nv), C.byref(na), C.byref(nf))
    vert = np.zeros(nv.value * 3)
    faces = np.zeros(na.value * 3, np.int32)
    fT, fw = np.zeros(nf.value), np.zeros(nf.value)
    rc = lib.dsmc_icosphere_build(levels, lon, body.t_night, body.t_day - body.t_night, body.mass, K_B,
                                  ptr(vert, C.c_double), ptr(faces, C.c_int32), ptr(fT, C.c_double),
                                  ptr(fw, C.c_double))
    if rc:
        raise RuntimeError(f"dsmc_icosphere_build: {rc}")
    return SurfaceArrays(levels=levels, vert=vert, faces=faces, face_T=fT, face_w=fw)


def body_params(lib, body: Body, n: int, lon: float, seed: int, hd: int = 40) -> Params:
    """The physical ballistic preset (dt 10 s, <= 20000 steps, photolysis lifetime
    6.4e5 s, freeze-out below 110 K, exponential shells alpha2 = 0.04 to 20 R)
    for `body`, with `hd` shells, `n` particles and the sun at `lon`."""
    p = preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=n, seed=seed)
    p.G = 1.0
    p.M = body.GM
    p.radius = body.radius
    p.mmass = body.mass
    p.mole = 1.0
    p.period = body.period
    p.v_esc = math.sqrt(2.0 * body.GM / body.radius)
    p.subsolar_lon = lon
    p.hd = hd
    p.max_height = 20.0 * body.radius
    p.unih = (p.max_height - p.radius) / ((1.0 + p.alpha2) ** hd - 1.0) * p.alpha2
    return p


def coll_params(lib, body: Body, n: int, lon: float, seed: int, hd: int = 40):
    """(phase 1, phase 2) of a physical collisional run of `body`: the ballistic
    preset keeping 64 speed bins to 2000 m/s per cell; emission Ceres' outgassing
    per unit area (2e26 molecules/s over R = 470 km) x10; phase 2 with dt 5 s,
    hard spheres d = 4.6e-10 m, degree-6 partner fits and the second random
    stream."""
    p1 = body_params(lib, body, n, lon, seed, hd)
    p1.production_rate = CERES_PRODUCTION * (body.radius / 470e3) ** 2
    p1.sublimation_scale = 10.0
    p1.vhist_bins = 64
    p1.vhist_max = 2000.0
    p1.fit_degree = 6
    p1.cross_section = math.pi * 4.6e-10 ** 2
    p2 = p1.copy()
    p2.dt = 5.0
    p2.stream = 2
    return p1, p2


@dataclass
class RunSpec:
    body: Body
    n: int
    lon: float
    seed: int
    collisional: bool = False


def batch_specs(nruns: int = 64, seed: int = 1, n_lo: float = 1e4, n_hi: float = 2e6,
                scale: float = 1.0) -> list[RunSpec]:
    """Runs of the batched workload: particle counts log-uniform on [n_lo, n_hi]
    (times `scale`), body uniform over BODIES, subsolar longitude U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nruns):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))) * scale))
        b = BODIES[int(rng.integers(len(BODIES)))]
        lon = float(rng.uniform(0.0, 2.0 * math.pi))
        out.append(RunSpec(b, max(n, 1), lon, 1000 + i))
    return out


def request_stream(rate: float, duration: float, seed: int = 5, n_lo: float = 1e4, n_hi: float = 4e6,
                   frac_collisional: float = 0.3, scale: float = 1.0) -> list[tuple[float, RunSpec]]:
    """Poisson arrivals at `rate` per second over `duration` seconds: (arrival
    time, run) pairs; counts log-uniform on [n_lo, n_hi] (times `scale`), a
    fraction of the requests flagged collisional."""
    rng = np.random.default_rng(seed)
    t, out, i = 0.0, [], 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi))) * scale))
        b = BODIES[int(rng.integers(len(BODIES)))]
        lon = float(rng.uniform(0.0, 2.0 * math.pi))
        coll = bool(rng.uniform() < frac_collisional)
        out.append((t, RunSpec(b, max(n, 1), lon, 100000 + i, coll)))
        i += 1
    return out


def build_runs(lib, specs: list[RunSpec], levels: int = 4, hd: int = 40):
    """(params, surfaces) of every run; surfaces of equal (body, lon) are shared."""
    params, surfs, cache = [], [], {}
    for s in specs:
        key = (s.body.name, s.lon)
        if key not in cache:
            cache[key] = icosphere(lib, levels, s.body, s.lon)
        surfs.append(cache[key])
        params.append(body_params(lib, s.body, s.n, s.lon, s.seed, hd))
    return params, surfs


def as_arrays(params: list[Params], surfs: list[SurfaceArrays]):
    """C arrays of dsmc_params and dsmc_surface for the batch entry points."""
    from .capi import Surface
    pa = (Params * len(params))(*params)
    sa = (Surface * len(surfs))(*[s.surface for s in surfs])
    return pa, sa
