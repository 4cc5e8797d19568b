"""ctypes mirror of include/dsmc_hip.h.

The structures here are field-for-field copies of the C structs; `load()` opens
the in-tree HIP library (dsmc_suite_amd/libdsmc_hip.so) and declares the
prototypes.  No compute happens in Python: every run goes through the C ABI.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libdsmc_hip.so"

MODE_COMPAT, MODE_PHYSICAL = 0, 1
STREAM_BALLISTIC, STREAM_ADVANCED = 1, 2
FATE_NONE, FATE_ESCAPED, FATE_FROZEN, FATE_PHOTOLYSED, FATE_DEPLETED, FATE_CAPPED = range(6)
PRESET_COMPAT_BALLISTIC, PRESET_COMPAT_ADVANCED = 0, 1
PRESET_CERES_PHYSICAL, PRESET_CERES_PHYSICAL_COLL = 2, 3
BATCH_PER_PARTICLE, BATCH_DENSITY_ONLY = 1, 2

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Params(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("stream", C.c_int), ("seed", C.c_uint64), ("nparticles", C.c_int64),
        ("G", C.c_double), ("M", C.c_double), ("radius", C.c_double),
        ("k_B", C.c_double), ("mmass", C.c_double), ("mole", C.c_double),
        ("lifetime", C.c_double), ("period", C.c_double), ("alpha", C.c_double),
        ("v_esc", C.c_double),
        ("dt", C.c_double), ("max_hop_steps", C.c_int64), ("max_steps", C.c_int64),
        ("t_freeze", C.c_double), ("w_min", C.c_double), ("subsolar_lon", C.c_double),
        ("ftd", C.c_int), ("ffd", C.c_int), ("hd", C.c_int), ("exp_grid", C.c_int),
        ("unih", C.c_double), ("alpha2", C.c_double), ("max_height", C.c_double),
        ("record_cap_cell", C.c_int), ("record_cap_total", C.c_int64),
        ("cross_section", C.c_double), ("fit_degree", C.c_int), ("max_per_cell", C.c_int),
        ("min_fit_samples", C.c_int), ("density_scale", C.c_double),
        ("sublimation_scale", C.c_double), ("production_rate", C.c_double),
        ("vhist_bins", C.c_int), ("vhist_max", C.c_double),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        return p


class Surface(C.Structure):
    _fields_ = [
        ("ftd", C.c_int), ("ffd", C.c_int), ("T_latlon", _dp),
        ("levels", C.c_int), ("nvert", C.c_int), ("vert", _dp), ("faces", _i32p),
        ("face_T", _dp), ("face_w", _dp),
    ]


class Result(C.Structure):
    _fields_ = [
        ("density", _dp), ("horizon", _dp), ("vmom", _dp), ("vcount", _i64p),
        ("ttime", _dp), ("end_state", _dp), ("steps", _i64p), ("fate", _i32p), ("hops", _i32p),
        ("vhist", _i64p),
        ("vrec", _dp), ("vrec_td", _i32p), ("vrec_fd", _i32p), ("vrec_ht", _i32p),
        ("vrec_n", C.c_int64),
        ("n_escaped", C.c_int64), ("n_frozen", C.c_int64), ("n_photolysed", C.c_int64),
        ("n_depleted", C.c_int64), ("n_capped", C.c_int64), ("n_collisions", C.c_int64),
        ("total_steps", C.c_int64), ("elapsed_ms", C.c_double),
    ]


class Phase2Input(C.Structure):
    _fields_ = [
        ("copy_atm", _dp), ("atm_ftd", C.c_int), ("atm_ffd", C.c_int), ("atm_hd", C.c_int),
        ("vrec", _dp), ("vrec_td", _i32p), ("vrec_fd", _i32p), ("vrec_ht", _i32p),
        ("nrec", C.c_int64), ("ndens", _dp), ("vhist", _i64p),
    ]


class LaunchShape(C.Structure):
    _fields_ = [("block", C.c_int), ("blocks_per_cu", C.c_int), ("relaunch_batch", C.c_int),
                ("step_budget", C.c_int), ("service_inline", C.c_int), ("late_step_budget", C.c_int),
                ("park_batch", C.c_int), ("late_inline", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("particle_steps", C.c_int64), ("queued", C.c_int64),
                ("running", C.c_int64), ("particles_in_flight", C.c_int64), ("slots", C.c_int64),
                ("rounds", C.c_int64), ("shape", LaunchShape)]


def ptr(a: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (or NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


class ThermalParams(C.Structure):
    """dsmc_thermal_params (the reference's ThermalParams, include/thermal.h)."""
    _fields_ = [("density", C.c_double), ("conductivity", C.c_double), ("Cp", C.c_double),
                ("albedo", C.c_double), ("distance_AU", C.c_double), ("depth", C.c_double),
                ("n_substep", C.c_int), ("z_angle", C.c_double), ("p_mat", C.c_double * 5),
                ("LH_mat", C.c_double * 5), ("tran_unit_p", C.c_double), ("tran_unit_LH", C.c_double)]


class SphGrid(C.Structure):
    """dsmc_sph_grid (the reference's SphericalGrid, include/types.h:56-64)."""
    _fields_ = [("ftd", C.c_int), ("ffd", C.c_int), ("hd", C.c_int), ("density", C.POINTER(C.c_double)),
                ("temperature", C.POINTER(C.c_double)), ("volume", C.POINTER(C.c_double)),
                ("horizon", C.POINTER(C.c_double))]


class Buffers:
    """Host arrays backing a Result; keeps them alive and exposes numpy views."""

    def __init__(self, p: Params, ttime_rows: int = 5, per_particle: bool = True,
                 vrec: bool = False, moments: bool = False, vhist: bool = False):
        n = int(p.nparticles)
        cells = p.hd * p.ftd * p.ffd
        self.density = np.zeros(cells)
        self.horizon = np.zeros(p.ftd * p.ffd)
        self.vmom = np.zeros(cells * 3) if moments else None
        self.vcount = np.zeros(cells, dtype=np.int64) if moments else None
        self.ttime = np.zeros(ttime_rows * n) if per_particle else None
        self.end_state = np.zeros(n * 6) if per_particle else None
        self.steps = np.zeros(n, dtype=np.int64) if per_particle else None
        self.fate = np.zeros(n, dtype=np.int32) if per_particle else None
        self.hops = np.zeros(n, dtype=np.int32) if per_particle else None
        self.vhist = np.zeros(cells * p.vhist_bins, dtype=np.int64) if vhist else None
        cap = int(p.record_cap_total) if vrec else 0
        self.vrec = np.zeros(cap * 3) if vrec else None
        self.vrec_td = np.zeros(cap, dtype=np.int32) if vrec else None
        self.vrec_fd = np.zeros(cap, dtype=np.int32) if vrec else None
        self.vrec_ht = np.zeros(cap, dtype=np.int32) if vrec else None
        r = Result()
        r.density = ptr(self.density, C.c_double)
        r.horizon = ptr(self.horizon, C.c_double)
        r.vmom = ptr(self.vmom, C.c_double)
        r.vcount = ptr(self.vcount, C.c_int64)
        r.ttime = ptr(self.ttime, C.c_double)
        r.end_state = ptr(self.end_state, C.c_double)
        r.steps = ptr(self.steps, C.c_int64)
        r.fate = ptr(self.fate, C.c_int32)
        r.hops = ptr(self.hops, C.c_int32)
        r.vhist = ptr(self.vhist, C.c_int64)
        r.vrec = ptr(self.vrec, C.c_double)
        r.vrec_td = ptr(self.vrec_td, C.c_int32)
        r.vrec_fd = ptr(self.vrec_fd, C.c_int32)
        r.vrec_ht = ptr(self.vrec_ht, C.c_int32)
        self.result = r

    def counters(self) -> dict:
        r = self.result
        return {k: int(getattr(r, k)) for k in
                ("n_escaped", "n_frozen", "n_photolysed", "n_depleted", "n_capped",
                 "n_collisions", "total_steps", "vrec_n")}


class SurfaceArrays:
    """Host arrays backing a Surface."""

    def __init__(self, T_latlon: np.ndarray | None = None, ftd: int = 0, ffd: int = 0,
                 levels: int = -1, vert=None, faces=None, face_T=None, face_w=None):
        s = Surface()
        self.T_latlon = None if T_latlon is None else np.ascontiguousarray(T_latlon, dtype=np.float64).ravel()
        s.ftd, s.ffd = ftd, ffd
        s.T_latlon = ptr(self.T_latlon, C.c_double)
        self.vert = None if vert is None else np.ascontiguousarray(vert, dtype=np.float64).ravel()
        self.faces = None if faces is None else np.ascontiguousarray(faces, dtype=np.int32).ravel()
        self.face_T = None if face_T is None else np.ascontiguousarray(face_T, dtype=np.float64)
        self.face_w = None if face_w is None else np.ascontiguousarray(face_w, dtype=np.float64)
        s.levels = levels
        s.nvert = 0 if self.vert is None else len(self.vert) // 3
        s.vert = ptr(self.vert, C.c_double)
        s.faces = ptr(self.faces, C.c_int32)
        s.face_T = ptr(self.face_T, C.c_double)
        s.face_w = ptr(self.face_w, C.c_double)
        self.surface = s


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Open the HIP library and declare its prototypes.  Raises if it is missing:
    there is no CPU fallback."""
    path = Path(path) if path else LIB_PATH
    if not path.exists():
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(str(path))
    P, S, R = C.POINTER(Params), C.POINTER(Surface), C.POINTER(Result)
    lib.dsmc_params_preset.argtypes = [P, C.c_int, C.c_int, C.c_int]
    lib.dsmc_run_ballistic.argtypes = [P, S, R]
    lib.dsmc_run_ballistic_noalloc.argtypes = [P, S, R]
    lib.dsmc_run_advanced.argtypes = [P, S, C.POINTER(Phase2Input), R]
    lib.dsmc_build_fits.argtypes = [P, C.POINTER(Phase2Input), _i32p, _dp, _dp, _dp, _i32p]
    lib.dsmc_run_batch.argtypes = [C.c_int, P, S, R, C.POINTER(LaunchShape)]
    lib.dsmc_number_density.argtypes = [P, S, _i64p, _dp]
    lib.dsmc_build_hist_fits.argtypes = [P, _i64p, _dp, _dp, _dp, _i32p]
    lib.dsmc_run_collisional.argtypes = [P, P, S, R, R]
    lib.dsmc_autotune.argtypes = [P, S, C.POINTER(LaunchShape)]
    lib.dsmc_icosphere_sizes.argtypes = [C.c_int, _i32p, _i32p, _i32p]
    lib.dsmc_icosphere_build.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, _dp, _i32p, _dp, _dp]
    lib.dsmc_icosphere_rasterise.argtypes = [S, C.c_int, C.c_int, _dp]
    lib.dsmc_batch_create.argtypes = [C.c_int, P, S, C.c_int, C.POINTER(LaunchShape), C.POINTER(C.c_int)]
    lib.dsmc_batch_create.restype = C.c_void_p
    lib.dsmc_batch_launch.argtypes = [C.c_void_p, C.POINTER(R)]
    lib.dsmc_batch_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), _dp, _i64p, C.POINTER(LaunchShape)]
    lib.dsmc_batch_destroy.argtypes = [C.c_void_p]
    lib.dsmc_batch_destroy.restype = None
    lib.dsmc_server_create.argtypes = [C.c_size_t, C.c_int, C.c_int]
    lib.dsmc_server_create.restype = C.c_void_p
    lib.dsmc_server_set_shape.argtypes = [C.c_void_p, C.POINTER(LaunchShape)]
    lib.dsmc_server_submit.argtypes = [C.c_void_p, P, S, R, C.POINTER(C.c_uint64)]
    lib.dsmc_server_submit_collisional.argtypes = [C.c_void_p, P, P, S, R, C.POINTER(C.c_uint64)]
    lib.dsmc_server_poll.argtypes = [C.c_void_p, C.c_uint64]
    lib.dsmc_server_wait.argtypes = [C.c_void_p, C.c_uint64]
    lib.dsmc_server_cancel.argtypes = [C.c_void_p, C.c_uint64]
    lib.dsmc_server_release.argtypes = [C.c_void_p, C.c_uint64]
    lib.dsmc_server_step.argtypes = [C.c_void_p]
    lib.dsmc_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
    lib.dsmc_server_destroy.argtypes = [C.c_void_p]
    lib.dsmc_server_destroy.restype = None
    lib.dsmc_run_create.argtypes = [P, S, R, C.POINTER(C.c_int)]
    lib.dsmc_run_create.restype = C.c_void_p
    lib.dsmc_run_launch.argtypes = [C.c_void_p, C.POINTER(LaunchShape)]
    lib.dsmc_run_sync.argtypes = [C.c_void_p]
    lib.dsmc_run_kernel_ms.argtypes = [C.c_void_p]
    lib.dsmc_run_kernel_ms.restype = C.c_double
    lib.dsmc_run_collect.argtypes = [C.c_void_p, R]
    lib.dsmc_run_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.dsmc_run_lean_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int64)]
    lib.dsmc_run_destroy.argtypes = [C.c_void_p]
    lib.dsmc_run_destroy.restype = None
    lib.dsmc_verlet_steps.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _dp]
    lib.dsmc_elementary.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int]
    lib.run_ballistic.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    lib.run_ballistic.restype = None
    lib.run_advanced.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    lib.run_advanced.restype = None
    lib.dsmc_version.restype = C.c_char_p
    lib.dsmc_device_info.argtypes = [C.c_char_p, C.c_size_t]
    return lib


def preset(lib: C.CDLL, which: int, ftd: int = 36, ffd: int = 72, n: int = 0, seed: int = 0) -> Params:
    p = Params()
    rc = lib.dsmc_params_preset(C.byref(p), which, ftd, ffd)
    if rc != 0:
        raise ValueError(f"dsmc_params_preset({which}) -> {rc}")
    p.nparticles = n
    p.seed = seed
    return p
