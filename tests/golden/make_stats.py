"""Reference statistics for tests/test_statistics.py.

Runs the COMPILED REFERENCE unmodified - glibc rand_r draws with its own
per-thread seeding - through oracle/_ref/ref_harness (built from the reference's
sources by `make -C oracle ref`), only where those sources exist, and stores what
its .dat outputs hold as a small fixture (data only):

  stats_reference.npz
    b_dims, b_ttime, b_coarse   compat ballistic, T_thermal.dat, 2000 particles,
                                one thread: the 5-row ttime and the density summed
                                over coarse cells (tests/helpers.coarse_density)
    a_dims, a_ttime, a_coarse   compat advanced, the phase-2 inputs of
                                advanced_keyed.npz (synthetic collisional density,
                                the reference's own phase-1 velocity records),
                                1500 particles, 8 threads: 9-row ttime (row 8 =
                                collisions per particle) and coarse density

    python tests/golden/make_stats.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
from tests.helpers import coarse_density  # noqa: E402

HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"
N_BALLISTIC, N_ADVANCED = 2000, 1500


def run(*args):
    subprocess.run([str(HARNESS), *map(str, args)], check=True, stdout=subprocess.DEVNULL)


def read_den(path: Path) -> tuple[np.ndarray, int, int, int]:
    v = np.loadtxt(path)
    ftd, ffd, hd = int(v[0]), int(v[1]), int(v[2])
    return v[3:], hd, ftd, ffd


def write_inputs(tmp: Path) -> tuple[Path, Path]:
    """The phase-2 inputs of advanced_keyed.npz as the reference reads them."""
    fx = np.load(HERE / "advanced_keyed.npz")
    ftd, ffd, hd = (int(x) for x in fx["in_den_dims"])
    den = tmp / "in_den.dat"
    with open(den, "w") as f:
        f.write(f"{ftd}\n{ffd}\n{hd}\n")
        f.write("".join("%9.6f\n" % v for v in fx["in_den"]))
    vrec = tmp / "in_vrec.dat"
    with open(vrec, "w") as f:
        f.write(f"{len(fx['in_vrec'])}\n")
        for v, t, d, h in zip(fx["in_vrec"], fx["in_vrec_td"], fx["in_vrec_fd"], fx["in_vrec_ht"]):
            f.write("%f %f %f\n%d\n%d\n%d\n" % (v[0], v[1], v[2], t, d, h))
    return den, vrec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    subprocess.run(["make", "-C", str(ROOT / "oracle"), f"REF={args.ref}", "ref"], check=True)
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        run("ballistic_randr", HERE / "T_thermal.dat", N_BALLISTIC, 1, tmp / "b")
        bden, hd, ftd, ffd = read_den(tmp / "b_den.dat")
        bt = np.loadtxt(tmp / "b_ttime.dat").reshape(5, N_BALLISTIC)
        den_in, vrec_in = write_inputs(tmp)
        run("advanced_randr", HERE / "T_thermal.dat", den_in, vrec_in, N_ADVANCED, 8, tmp / "a")
        aden, ahd, aftd, affd = read_den(tmp / "a_den.dat")
        at = np.loadtxt(tmp / "a_ttime.dat").reshape(9, N_ADVANCED)
    np.savez_compressed(HERE / "stats_reference.npz",
                        b_dims=np.array([hd, ftd, ffd, N_BALLISTIC]), b_ttime=bt,
                        b_coarse=coarse_density(bden, hd, ftd, ffd),
                        a_dims=np.array([ahd, aftd, affd, N_ADVANCED]), a_ttime=at,
                        a_coarse=coarse_density(aden, ahd, aftd, affd, ht_edges=(0, 1, 2, 3, 4, 8, 16, 50)))
    f = HERE / "stats_reference.npz"
    print(f"{f.name}: {f.stat().st_size} B; ballistic escaped {int((bt[1:] == 0).all(axis=0).sum())}, "
          f"advanced collisions {int(at[8].sum())}")


if __name__ == "__main__":
    main()
