#!/usr/bin/env python3
"""Generate the golden fixtures of tests/golden/ from the COMPILED REFERENCE.

Runs only where the dsmc_suite sources are present (default /root/reference):
it builds oracle/_ref/ref_harness with `make -C oracle ref` (the reference's own
hot-path sources compiled as they lie, linked with --wrap to inject Philox draws
and capture full-precision outputs) and writes small .npz fixtures.  The fixtures
are data only: inputs and the reference's outputs.

    python tests/golden/make_golden.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import hashlib
import os
import struct
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"


def read_bin(path: Path) -> dict:
    out, b, o = {}, path.read_bytes(), 0
    while o < len(b):
        name = b[o:o + 32].split(b"\0")[0].decode()
        kind = chr(b[o + 32])
        es, cnt = struct.unpack("<qq", b[o + 33:o + 49])
        o += 49
        dt = {("f", 8): np.float64, ("i", 8): np.int64, ("i", 4): np.int32}[(kind, es)]
        out[name] = np.frombuffer(b[o:o + es * cnt], dtype=dt).copy()
        o += es * cnt
    return out


def sparse(a: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    idx = np.nonzero(a)[0].astype(np.int32)
    return idx, a[idx]


def run(*args):
    subprocess.run([str(HARNESS), *map(str, args)], check=True)


def write_den_dat(path: Path, den: np.ndarray, ftd: int, ffd: int, hd: int):
    with open(path, "w") as f:
        f.write(f"{ftd}\n{ffd}\n{hd}\n")
        f.write("".join("%9.6f\n" % v for v in den))


def synthetic_T(ftd=36, ffd=72) -> str:
    """A second T field with a cold polar cap (exercises permanent freezing)."""
    lat = (np.arange(ftd) + 0.5) / ftd * np.pi - np.pi / 2
    lon = (np.arange(ffd) + 0.5) / ffd * 2 * np.pi
    cz = np.cos(lat)[:, None] * np.cos(lon[None, :] - np.pi / 2)
    T = 80 + 160 * np.maximum(cz, 0) ** 0.25
    T[np.abs(lat) > np.radians(80)] = 90.0
    return f"{ftd}\n{ffd}\n" + "".join("%8.6f\n" % v for v in T.ravel())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    if not Path(args.ref, "src", "dsmc", "ballistic.c").exists():
        sys.exit(f"reference sources not found under {args.ref}")
    subprocess.run(["make", "-C", str(ROOT / "oracle"), f"REF={args.ref}", "ref"], check=True)
    tmp = Path(tempfile.mkdtemp(prefix="dsmc_golden_"))

    # ---- inputs: the reference thermal model's T field, and a synthetic one
    run("thermal", tmp / "T_thermal.dat")
    (HERE / "T_thermal.dat").write_text((tmp / "T_thermal.dat").read_text())
    (HERE / "T_polar.dat").write_text(synthetic_T())

    # ---- velocity table + test_verlet.c:76-86 known answer (glibc rand_r, seed 42)
    run("vtable", tmp / "vt.bin")
    vt = read_bin(tmp / "vt.bin")
    np.savez_compressed(HERE / "vtable.npz", temps=vt["temps"], speeds=vt["speeds"].reshape(23, 1000),
                        count=vt["count"], mean200=vt["mean200"])

    # ---- polyfit: test_verlet.c:99-135 cubic + random degree-3 and degree-6 sets
    rng = np.random.default_rng(11)
    sets = [(np.arange(11.0), 2 * np.arange(11.0) ** 3 - np.arange(11.0) ** 2 + 3 * np.arange(11.0) - 1, 3)]
    for k in range(40):
        n = int(rng.integers(4, 40))
        deg = 3 if k < 25 else int(rng.integers(1, 7))
        x = np.sort(rng.uniform(0, 150, n))
        y = np.cumsum(rng.integers(0, 5, n))[::-1].astype(float)
        sets.append((x, y, deg))
    with open(tmp / "pf_in.bin", "wb") as f:
        f.write(struct.pack("<q", len(sets)))
        for x, y, d in sets:
            f.write(struct.pack("<qq", len(x), d))
            f.write(np.asarray(x, np.float64).tobytes())
            f.write(np.asarray(y, np.float64).tobytes())
    run("polyfit", tmp / "pf_in.bin", tmp / "pf_out.bin")
    raw = np.frombuffer((tmp / "pf_out.bin").read_bytes(), dtype=np.float64).reshape(len(sets), 8)
    np.savez_compressed(HERE / "polyfit.npz",
                        x=np.concatenate([s[0] for s in sets]), y=np.concatenate([s[1] for s in sets]),
                        n=np.array([len(s[0]) for s in sets]), deg=np.array([s[2] for s in sets]),
                        rc=raw[:, 0].view(np.int64), coeffs=raw[:, 1:])

    # ---- ballistic, keyed Philox streams, one particle per reference call
    for tname, N, seed in (("T_thermal", 200, 7), ("T_polar", 60, 3)):
        out = tmp / f"bk_{tname}.bin"
        run("ballistic_keyed", HERE / f"{tname}.dat", N, seed, out, 3, 4000)
        g = read_bin(out)
        di, dv = sparse(g["density"])
        hi, hv = sparse(g["horizon"])
        tr = np.frombuffer((tmp / f"bk_{tname}.bin.trace").read_bytes(), dtype=np.float64).reshape(-1, 7)
        np.savez_compressed(HERE / f"ballistic_keyed_{tname}.npz", dims=g["dims"], seed=seed,
                            ttime=g["ttime"].reshape(5, N), steps=g["steps"], draws=g["draws"],
                            end_state=g["end_state"].reshape(N, 6), den_idx=di, den_val=dv,
                            hor_idx=hi, hor_val=hv, trace_pid=3, trace=tr)

    # ---- ballistic, one call with a sequential stream: v_record cap semantics
    out = tmp / "bseq.bin"
    run("ballistic_seq", HERE / "T_thermal.dat", 120, 5, out)
    g = read_bin(out)
    di, dv = sparse(g["density"])
    hi, hv = sparse(g["horizon"])
    np.savez_compressed(HERE / "ballistic_seq.npz", dims=g["dims"], seed=5, ttime=g["ttime"].reshape(5, -1),
                        den_idx=di, den_val=dv, hor_idx=hi, hor_val=hv, vrec=g["vrec"].reshape(-1, 3),
                        vrec_td=g["vrec_td"], vrec_fd=g["vrec_fd"], vrec_ht=g["vrec_ht"], totals=g["totals"])

    # ---- unmodified reference (glibc rand_r, 1 thread): text outputs, hashed
    run("ballistic_randr", HERE / "T_thermal.dat", 25, 1, tmp / "rr")
    hashes = {}
    for suf in ("den", "horizon", "v_record", "ttime"):
        hashes[suf] = hashlib.sha256((tmp / f"rr_{suf}.dat").read_bytes()).hexdigest()
    (HERE / "ballistic_randr_sha256.txt").write_text(
        "# ref_harness ballistic_randr T_thermal.dat 25 particles, 1 thread\n" +
        "".join(f"{k} {v}\n" for k, v in hashes.items()))

    # ---- advanced: phase-1 records from a reference run, synthetic collisional density
    run("ballistic_randr", HERE / "T_thermal.dat", 150, 1, tmp / "p1")
    lines = (tmp / "p1_v_record.dat").read_text().split("\n")
    nrec = int(lines[0])
    recs = np.array([list(map(float, lines[1 + 4 * i].split())) for i in range(nrec)])
    itd = np.array([int(lines[2 + 4 * i]) for i in range(nrec)], np.int32)
    ifd = np.array([int(lines[3 + 4 * i]) for i in range(nrec)], np.int32)
    iht = np.array([int(lines[4 + 4 * i]) for i in range(nrec)], np.int32)
    ftd, ffd, hd1 = 36, 72, 100
    ht = np.arange(hd1)[:, None, None]
    lat = ((np.arange(ftd) + 0.5) / ftd * np.pi - np.pi / 2)[None, :, None]
    den = (3.0e15 * np.exp(-ht / 4.0) * (0.3 + np.cos(lat) ** 2) * np.ones((1, 1, ffd))).ravel()
    write_den_dat(tmp / "p1s_den.dat", den, ftd, ffd, hd1)
    den_parsed = np.loadtxt(tmp / "p1s_den.dat", skiprows=3)
    out = tmp / "ak.bin"
    N = 24
    run("advanced_keyed", HERE / "T_thermal.dat", tmp / "p1s_den.dat", tmp / "p1_v_record.dat", N, 9, out)
    g = read_bin(out)
    di, dv = sparse(g["density"])
    hi, hv = sparse(g["horizon"])
    np.savez_compressed(HERE / "advanced_keyed.npz", dims=g["dims"], seed=9, ttime=g["ttime"].reshape(9, N),
                        steps=g["steps"], draws=g["draws"], end_state=g["end_state"].reshape(N, 6),
                        den_idx=di, den_val=dv, hor_idx=hi, hor_val=hv,
                        in_den=den_parsed, in_den_dims=np.array([ftd, ffd, hd1]),
                        in_vrec=recs, in_vrec_td=itd, in_vrec_fd=ifd, in_vrec_ht=iht)
    for f in sorted(HERE.glob("*")):
        if f.suffix in (".npz", ".dat", ".txt"):
            print(f"{f.name:40s} {f.stat().st_size:>9d} B")


if __name__ == "__main__":
    main()
