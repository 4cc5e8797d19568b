"""CPU-side checks of the product library and its host logic (no GPU calls).

* libdsmc_hip.so loads and exports every function include/*.h declares;
* presets reproduce the reference's constants;
* the per-lane state machine the kernels run (dsmc_lane.h, built for the host by
  tests/emu/lane_emu.cpp) matches the oracle's device-math build bit for bit;
* the icosphere builder, the .dat readers/writers and the device-math functions.
"""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from dsmc_suite_amd import capi
from dsmc_suite_amd.capi import (PRESET_CERES_PHYSICAL, PRESET_COMPAT_ADVANCED, PRESET_COMPAT_BALLISTIC,
                                 Buffers, SurfaceArrays, ptr)
from tests.helpers import GOLDEN, ROOT, compat_params, oracle, read_T, run_oracle


@pytest.fixture(scope="module")
def lib():
    return capi.load()


@pytest.fixture(scope="module")
def emu():
    so = ROOT / "tests" / "emu" / "liblane_emu.so"
    src = ROOT / "tests" / "emu" / "lane_emu.cpp"
    if not so.exists() or so.stat().st_mtime < max(p.stat().st_mtime for p in
                                                   [src, *(ROOT / "dsmc_suite_amd" / "csrc").glob("*.h")]):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(so),
                        str(src)], check=True)
    return C.CDLL(str(so))


def declared_functions():
    text = "".join(h.read_text() for h in sorted((ROOT / "include").glob("*.h")))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M):
        name = m.group(1)
        if name not in ("if", "return", "sizeof"):
            names.add(name)
    return sorted(names)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert "run_ballistic" in names and "write_density_dat" in names and "dsmc_run_create" in names
    assert "verlet_step" in names and "polyfit" in names and "grid_compute_volume_exp" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_grid_create_and_free(lib):
    """grid.c:7-26: the four arrays zeroed, sized [hd*ftd*ffd] / [ftd*ffd]."""
    lib.grid_create.restype = C.POINTER(capi.SphGrid)
    lib.grid_create.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.grid_free.argtypes = [C.POINTER(capi.SphGrid)]
    gp = lib.grid_create(36, 72, 5)
    g = gp.contents
    assert (g.ftd, g.ffd, g.hd) == (36, 72, 5)
    for name, n in (("density", 5 * 36 * 72), ("volume", 5 * 36 * 72), ("temperature", 36 * 72),
                    ("horizon", 36 * 72)):
        a = np.ctypeslib.as_array(getattr(g, name), shape=(n,))
        assert not a.any(), name
        a[:] = 1.0                                     # the whole extent is writable
    lib.grid_free(gp)
    assert not lib.grid_create(-1, 72, 5) and not lib.grid_create(2 ** 30, 2 ** 30, 2 ** 30)
    lib.grid_free(None)


def test_reference_cli_links_against_the_library():
    """The reference CLI (src/main.c, unmodified; make -C oracle ref) linked
    against libdsmc_hip.so alone: with every symbol bound at load time it
    starts and prints its usage (no command: no device work)."""
    cli = ROOT / "oracle" / "_ref" / "dsmc_cli"
    if not cli.exists():
        pytest.skip("the reference CLI (make -C oracle ref) is not built")
    ldd = subprocess.run(["readelf", "-d", str(cli)], capture_output=True, text=True, check=True).stdout
    assert "[libdsmc_hip.so]" in ldd
    r = subprocess.run([str(cli)], capture_output=True, text=True, timeout=60,
                       env={"LD_BIND_NOW": "1", "PATH": "/usr/bin:/bin"})
    assert r.returncode == 1 and "pipeline  - Run full pipeline sequentially" in r.stdout, r.stderr


def test_compat_presets_equal_reference_constants(lib):
    for which, adv in ((PRESET_COMPAT_BALLISTIC, False), (PRESET_COMPAT_ADVANCED, True)):
        got = capi.preset(lib, which, 36, 72, n=5, seed=3)
        want = compat_params(36, 72, 5, 3, advanced=adv)
        for f, _ in capi.Params._fields_:
            if f in ("sublimation_scale",):
                continue
            assert getattr(got, f) == getattr(want, f), f


def build_icosphere(lib, levels):
    nv, na, nf = C.c_int32(), C.c_int32(), C.c_int32()
    assert lib.dsmc_icosphere_sizes(levels, C.byref(nv), C.byref(na), C.byref(nf)) == 0
    vert = np.zeros(nv.value * 3)
    faces = np.zeros(na.value * 3, np.int32)
    fT, fw = np.zeros(nf.value), np.zeros(nf.value)
    assert lib.dsmc_icosphere_build(levels, 0.0, 100.0, 135.0, 2.99e-26, 1.380649e-23, ptr(vert, C.c_double),
                                    ptr(faces, C.c_int32), ptr(fT, C.c_double), ptr(fw, C.c_double)) == 0
    return vert.reshape(-1, 3), faces.reshape(-1, 3), fT, fw


def reference_icosphere(levels):
    """scripts/vesta_raytrace_gif.py:35-102 restated with numpy (vertex order)."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(levels):
        nv, mid, nf = list(v), {}, []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = (v[a] + v[b]) / 2
                nv.append(p / np.linalg.norm(p))
                mid[k] = len(nv) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        v, f = np.array(nv), nf
    return v, np.array(f)


@pytest.mark.parametrize("levels", [0, 3, 4])
def test_icosphere(lib, levels):
    vert, faces, fT, fw = build_icosphere(lib, levels)
    nfine = 20 * 4 ** levels
    assert len(vert) == 10 * 4 ** levels + 2 and len(faces) == sum(20 * 4 ** l for l in range(levels + 1))
    rv, rf = reference_icosphere(levels)
    assert np.allclose(vert, rv, rtol=0, atol=1e-15)
    fine = faces[-nfine:]
    assert np.array_equal(fine, rf)
    a, b, c = vert[fine[:, 0]], vert[fine[:, 1]], vert[fine[:, 2]]
    cen = a + b + c
    assert np.all(np.einsum("ij,ij->i", np.cross(b - a, c - a), cen) > 0)          # outward, consistent
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()      # flat faces: below 4 pi,
    assert area < 4 * np.pi and (levels < 3 or area > 0.97 * 4 * np.pi)     # converging with level
    cz = cen[:, 1] / np.linalg.norm(cen, axis=1)                                    # subsolar (lat 0, lon 0) = +y
    assert np.allclose(fT, 100 + 135 * np.maximum(cz, 0) ** 0.25, rtol=1e-14)
    flux = fw / (0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1))   # Antoine flux, monotone in T
    assert np.all(fw >= 0) and np.all(np.diff(flux[np.argsort(fT)]) >= -1e-12 * flux.max())


@pytest.mark.parametrize("levels", [0, 3, 4, 5])
def test_face_map_locates_as_the_full_search(lib, emu, levels):
    """locate_face_near (hint, then the cube-map candidates, then the search) gives
    the hierarchical search's face for random directions and for the hard ones:
    vertices, edge midpoints, points 1e-13..1e-9 off every kind of edge, cube
    edges and face-map cell boundaries, zero and non-finite directions."""
    vert, faces, fT, fw = build_icosphere(lib, levels)
    nfine = 20 * 4 ** levels
    fine = faces[-nfine:]
    rng = np.random.default_rng(levels)
    R = 470e3
    d = [rng.normal(size=(20000, 3)) * R, vert * R]
    a, b, c = vert[fine[:, 0]], vert[fine[:, 1]], vert[fine[:, 2]]
    d += [(a + b + c) * R, (a + b) * R / 2]
    for p0, p1 in ((a, b), (b, c), (c, a)):
        t = rng.uniform(size=(nfine, 1))
        on = p0 + t * (p1 - p0)
        nrm = np.cross(p0, p1)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        for delta in (1e-13, 1e-12, 1e-11, 1e-9):
            d += [(on + delta * nrm) * R, (on - delta * nrm) * R]
    G = int(min(128, max(8, round((nfine / 1.25) ** 0.5))))
    k = np.arange(G + 1) * 2.0 / G - 1.0
    u, v = np.meshgrid(k, rng.uniform(-1, 1, 7))
    for col in range(3):                                    # cell boundaries on each cube face
        for s in (1.0, -1.0):
            q = np.zeros((u.size, 3))
            q[:, col] = s
            q[:, (col + 1) % 3] = u.ravel()
            q[:, (col + 2) % 3] = v.ravel()
            d += [q * R, q[:, [0, 2, 1]] * R]
    w = rng.uniform(-1, 1, 500)
    d += [np.stack([np.ones_like(w), np.ones_like(w), w], 1), np.stack([w, -np.ones_like(w), np.ones_like(w)], 1)]
    d += [np.array([[0.0, 0.0, 0.0], [np.inf, 1.0, 0.0], [1e300, 1e300, 1.0], [1e-300, 0.0, 0.0]])]
    dirs = np.ascontiguousarray(np.concatenate(d))
    n = len(dirs)
    hints = rng.integers(-1, nfine, n).astype(np.int32)
    surf = SurfaceArrays(levels=levels, vert=vert.ravel(), faces=faces.ravel(), face_T=fT, face_w=fw)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=1, seed=1)
    near, full = np.zeros(n, np.int32), np.zeros(n, np.int32)
    emu.emu_locate.argtypes = [C.c_void_p, C.c_void_p, C.c_int64] + [C.c_void_p] * 4
    for h in (None, hints):
        assert emu.emu_locate(C.byref(p), C.byref(surf.surface), n, dirs.ctypes.data,
                              None if h is None else h.ctypes.data, near.ctypes.data, full.ctypes.data) == 0
        bad = np.nonzero(near != full)[0]
        assert len(bad) == 0, (len(bad), dirs[bad[:5]], near[bad[:5]], full[bad[:5]])
    assert len(np.unique(full)) == nfine
    # the pools' search: the first candidates' normals stored in the face-map cell
    pool = np.zeros(n, np.int32)
    emu.emu_locate_pool.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    assert emu.emu_locate_pool(C.byref(p), C.byref(surf.surface), n, dirs.ctypes.data, pool.ctypes.data) == 0
    bad = np.nonzero(pool != full)[0]
    assert len(bad) == 0, (len(bad), dirs[bad[:5]], pool[bad[:5]], full[bad[:5]])


@pytest.mark.parametrize("levels", [3, 4])
def test_face_map_candidates_cover_the_sphere(lib, emu, levels):
    """The face map's first four candidates hold nearly every direction firmly, so
    the landing service's face search rarely needs the second group and almost
    never the hierarchical search (one lane that does holds up the wavefront)."""
    vert, faces, fT, fw = build_icosphere(lib, levels)
    surf = SurfaceArrays(levels=levels, vert=vert.ravel(), faces=faces.ravel(), face_T=fT, face_w=fw)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=1, seed=1)
    d = np.ascontiguousarray(np.random.default_rng(7).normal(size=(200000, 3)))
    cnt = np.zeros(3, np.int64)
    emu.emu_locate_paths.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    assert emu.emu_locate_paths(C.byref(p), C.byref(surf.surface), len(d), d.ctypes.data, cnt.ctypes.data) == 0
    first, second, search = cnt / len(d)
    assert first > 0.98 and second < 0.02 and search < 1e-3, (first, second, search)


@pytest.mark.parametrize("tname", ["T_thermal", "T_polar"])
def test_lane_state_machine_matches_oracle_compat(emu, tname):
    dev = oracle(devmath=True)
    T, ftd, ffd = read_T(GOLDEN / f"{tname}.dat")
    p = compat_params(ftd, ffd, 150, 21)
    surf = SurfaceArrays(T, ftd, ffd)
    want = run_oracle(dev, p, surf, nthreads=4)
    got = Buffers(p, 5)
    assert emu.emu_ballistic(C.byref(p), C.byref(surf.surface), C.byref(got.result)) == 0
    for k in ("ttime", "steps", "end_state", "fate", "hops"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    assert np.allclose(got.density, want.density, rtol=1e-12, atol=1e-12)


def test_lane_state_machine_matches_oracle_physical(lib, emu):
    dev = oracle(devmath=True)
    vert, faces, fT, fw = build_icosphere(lib, 3)
    surf = SurfaceArrays(levels=3, vert=vert, faces=faces, face_T=fT, face_w=fw)
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=600, seed=9)
    want = run_oracle(dev, p, surf, nthreads=4, moments=True)
    got = Buffers(p, 5, moments=True)
    assert emu.emu_ballistic(C.byref(p), C.byref(surf.surface), C.byref(got.result)) == 0
    for k in ("steps", "end_state", "fate", "hops", "vcount"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    assert np.allclose(got.density, want.density, rtol=1e-12, atol=0)


@pytest.mark.parametrize("body,lifetime", [(0, None), (3, None), (0, 1e30)])
@pytest.mark.parametrize("budget,path", [(23, 3), (4096, 3), (23, 4), (300, 4)])
def test_lean_flight_path_matches_oracle(lib, emu, body, lifetime, budget, path):
    """The event-batched physical flight kernel's lane sequence (lean steps
    while nothing but the step count, the sunlit count and the open visit
    change, the full step logic when one needs it, the cell entries reloaded
    after it; records and the service kernel between launches): bit-identical
    to the oracle for a long-hop body (Ceres) and a short-hop one (Mercury),
    with velocity moments; and with a photolysis lifetime so long that the
    lean steps' sunlit-count limit is clamped to INT32_MAX.  Path 4: two such
    launches, then the kernel with the landing service inline takes each
    particle over from the record they left (late_inline)."""
    from dsmc_suite_amd import workloads as W
    dev = oracle(devmath=True)
    b = W.BODIES[body]
    surf = W.icosphere(lib, 3, b, 2.0)
    p = W.body_params(lib, b, 150, 2.0, 91, hd=40)
    if lifetime:
        p.lifetime = lifetime
    want = run_oracle(dev, p, surf, nthreads=8, moments=True)
    got = Buffers(p, 5, moments=True)
    emu.emu_ballistic_path.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int64]
    assert emu.emu_ballistic_path(C.byref(p), C.byref(surf.surface), C.byref(got.result), path, budget) == 0
    for k in ("steps", "end_state", "fate", "hops", "vcount", "horizon"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    cnt = np.repeat(want.vcount.astype(np.float64), 3)
    assert np.all(np.abs(got.vmom - want.vmom) <= 1e-12 * cnt * 5e3 + 1e-9)


@pytest.mark.parametrize("path,budget", [(1, 37), (2, 37), (1, 4096), (2, 1024)])
def test_record_paths_match_oracle(lib, emu, path, budget):
    """The kernels' record traffic (split flight/service kernels, or the physical
    flight kernel with the service inline), each launch starting from garbage
    registers, with short step budgets (many pauses): bit-identical to the oracle
    for a short-hop body (many landings) and for compat runs."""
    from dsmc_suite_amd import workloads as W
    dev = oracle(devmath=True)
    b = W.BODIES[3]
    surf = W.icosphere(lib, 3, b, 2.0)
    p = W.body_params(lib, b, 60, 2.0, 77, hd=40)
    want = run_oracle(dev, p, surf, nthreads=4, moments=True)
    got = Buffers(p, 5, moments=True)
    emu.emu_ballistic_path.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int64]
    assert emu.emu_ballistic_path(C.byref(p), C.byref(surf.surface), C.byref(got.result), path, budget) == 0
    for k in ("steps", "end_state", "fate", "hops", "vcount", "horizon"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    assert want.hops.mean() > 20
    # compat runs on both kernel paths: split, and the flight loop with the service inline
    T, ftd, ffd = read_T(GOLDEN / "T_thermal.dat")
    p = compat_params(ftd, ffd, 80, 5)
    s = SurfaceArrays(T, ftd, ffd)
    want = run_oracle(dev, p, s, nthreads=4)
    got = Buffers(p, 5)
    assert emu.emu_ballistic_path(C.byref(p), C.byref(s.surface), C.byref(got.result), path, budget) == 0
    for k in ("ttime", "steps", "end_state", "fate", "hops"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    assert np.allclose(got.density, want.density, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("body", range(4))
def test_lane_state_machine_matches_oracle_bodies(lib, emu, body):
    """Every body of the multi-run workloads, on the lane state machine: short
    hops and many landings (the guessed-face test and the cell walks from the
    landing cell must give what the oracle's full searches give)."""
    from dsmc_suite_amd import workloads as W
    b = W.BODIES[body]
    dev = oracle(devmath=True)
    surf = W.icosphere(lib, 3, b, 0.7 + body)
    p = W.body_params(lib, b, 120, 0.7 + body, 40 + body, hd=40)
    want = run_oracle(dev, p, surf, nthreads=4, moments=True)
    got = Buffers(p, 5, moments=True)
    assert emu.emu_ballistic(C.byref(p), C.byref(surf.surface), C.byref(got.result)) == 0
    for k in ("steps", "end_state", "fate", "hops", "vcount", "horizon"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters()
    assert want.hops.sum() > len(want.hops)          # relaunches happened


def test_device_math_accuracy_against_glibc():
    """dsmc_math.h against glibc: at most 1 ulp for sin/cos/exp/log and 3 ulp for
    asin/atan2 on the argument ranges the kernels use."""
    libm, dev = oracle(False), oracle(True)
    rng = np.random.default_rng(1)
    n = 300000
    cases = {0: (rng.uniform(-12, 12, n), None, 1), 1: (rng.uniform(-12, 12, n), None, 1),
             2: (rng.uniform(-1, 1, n), None, 3), 3: (rng.uniform(-1, 1, n) * 5e5, rng.uniform(-1, 1, n) * 5e5, 3),
             4: (rng.uniform(-745, 709, n), None, 1), 5: (np.exp(rng.uniform(-700, 700, n)), None, 1)}
    for fn, (x, y, bound) in cases.items():
        yy = x if y is None else y
        a, b = np.zeros(n), np.zeros(n)
        libm.cpu_ref_elementary(fn, ptr(x, C.c_double), ptr(yy, C.c_double), ptr(a, C.c_double), n)
        dev.cpu_ref_elementary(fn, ptr(x, C.c_double), ptr(yy, C.c_double), ptr(b, C.c_double), n)
        ok = np.isfinite(a) & (a != 0)
        ulp = np.abs(a[ok] - b[ok]) / np.spacing(np.abs(a[ok]))
        assert ulp.max() <= bound, (fn, ulp.max())
    # special values
    x = np.array([np.pi, np.pi / 2, 0.0, -0.0, 1.0, -1.0])
    out = np.zeros(6)
    dev.cpu_ref_elementary(0, ptr(x, C.c_double), ptr(x, C.c_double), ptr(out, C.c_double), 6)
    assert out[0] == np.sin(np.pi) and out[2] == 0.0


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    lib = oracle(False)
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        lib.cpu_ref_philox(c, k, o)
        assert tuple(o) == want


def test_dat_roundtrip_and_errors(lib, tmp_path):
    lib.read_temperature_dat.restype = C.POINTER(C.c_double)
    lib.read_temperature_dat.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.read_density_dat.restype = C.POINTER(C.c_double)
    lib.read_density_dat.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
    lib.write_density_dat.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int]
    ftd, ffd = C.c_int(), C.c_int()
    T = lib.read_temperature_dat(str(GOLDEN / "T_thermal.dat").encode(), C.byref(ftd), C.byref(ffd))
    assert (ftd.value, ffd.value) == (36, 72)
    ref, _, _ = read_T(GOLDEN / "T_thermal.dat")
    assert np.array_equal(np.ctypeslib.as_array(T, (36 * 72,)), ref)
    den = np.random.default_rng(0).uniform(0, 1e4, 2 * 3 * 4)
    f = str(tmp_path / "d.dat").encode()
    assert lib.write_density_dat(f, ptr(den, C.c_double), 3, 4, 2) == 0
    text = (tmp_path / "d.dat").read_text()
    assert text == "3\n4\n2\n" + "".join("%9.6f\n" % v for v in den)        # fileio.c:27-45
    hd = C.c_int()
    back = lib.read_density_dat(f, C.byref(ftd), C.byref(ffd), C.byref(hd), 1)
    assert np.array_equal(np.ctypeslib.as_array(back, (24,)), np.array(["%9.6f" % v for v in den], float))
    # missing and short files fail loudly instead of leaving zeros
    assert not lib.read_temperature_dat(str(tmp_path / "nope.dat").encode(), C.byref(ftd), C.byref(ffd))
    (tmp_path / "short.dat").write_text("36\n72\n1.0\n2.0\n")
    assert not lib.read_temperature_dat(str(tmp_path / "short.dat").encode(), C.byref(ftd), C.byref(ffd))


def test_invalid_arguments_rejected_without_device(lib):
    p = capi.preset(lib, PRESET_CERES_PHYSICAL, 36, 72, n=10)
    s = SurfaceArrays()           # physical run without a surface
    b = Buffers(p, 5)
    assert lib.dsmc_params_preset(C.byref(p), 99, 36, 72) == capi.C.c_int(-1).value
    assert lib.dsmc_icosphere_sizes(11, None, None, None) == -1


def oracle_fits(dev, p, inp):
    ncell = p.ftd * p.ffd * p.hd
    cnt = np.zeros(ncell, np.int32)
    sp = np.zeros(ncell * p.max_per_cell)
    dev.cpu_ref_bin_records(inp.vrec, inp.vrec_td, inp.vrec_fd, inp.vrec_ht, C.c_int64(inp.nrec), p.ftd, p.ffd,
                            p.hd, p.max_per_cell, ptr(cnt, C.c_int32), ptr(sp, C.c_double))
    co = np.zeros(ncell * (p.fit_degree + 1))
    mn, mx, va = np.zeros(ncell), np.zeros(ncell), np.zeros(ncell, np.int32)
    dev.cpu_ref_build_fits(ptr(cnt, C.c_int32), ptr(sp, C.c_double), p.ftd, p.ffd, p.hd, p.max_per_cell,
                           p.fit_degree, p.min_fit_samples, ptr(co, C.c_double), ptr(mn, C.c_double),
                           ptr(mx, C.c_double), ptr(va, C.c_int32), 4)
    return cnt, co, mn, mx, va


def test_lane_state_machine_matches_oracle_advanced(emu):
    """Collisional phase (advanced.c:241-450), fits supplied from the oracle."""
    from tests.helpers import golden, phase2
    dev = oracle(devmath=True)
    fx = golden("advanced_keyed.npz")
    T, ftd, ffd = read_T(GOLDEN / "T_thermal.dat")
    p = compat_params(ftd, ffd, 24, 9, advanced=True)
    inp, keep = phase2(fx)
    surf = SurfaceArrays(T, ftd, ffd)
    want = run_oracle(dev, p, surf, rows=9, phase2_input=inp, nthreads=8)
    cnt, co, mn, mx, va = oracle_fits(dev, p, inp)
    got = Buffers(p, 9)
    assert emu.emu_advanced(C.byref(p), C.byref(surf.surface), inp.copy_atm, C.c_int64(len(keep[0])),
                            ptr(cnt, C.c_int32), ptr(co, C.c_double), ptr(mn, C.c_double), ptr(mx, C.c_double),
                            ptr(va, C.c_int32), C.byref(got.result)) == 0
    for k in ("ttime", "steps", "end_state", "fate", "hops"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    assert got.counters() == want.counters() and want.result.n_collisions > 100

