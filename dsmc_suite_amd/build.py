"""Build recipe of libdsmc_hip.so (hipcc, gfx950), in-tree.

    python -m dsmc_suite_amd.build        # or __graft_entry__.build()

Every HIP translation unit is compiled with -ffp-contract=off (the kernels'
arithmetic must match the host bit for bit), -munsafe-fp-atomics (hardware FP64
atomics on coarse-grained device memory) and without packed FP32 operations.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libdsmc_hip.so"
ARCH = os.environ.get("DSMC_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "host.hip", "advanced.hip", "pool.hip", "batch.hip", "serve.hip", "thermal.hip",
               "shadow.hip"]
CXX_SOURCES = ["io.cpp", "surface.cpp", "common_api.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _compile(args):
    cmd, src = args
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{' '.join(cmd)}\n{err}")
    return err


def build(verbose: bool = False, oracle: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    cc = hipcc()
    jobs, objs = [], []
    for s in HIP_SOURCES:
        o = BUILD / (s + ".o")
        jobs.append(([cc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / s), "-o", str(o)], s))
        objs.append(o)
    for s in CXX_SOURCES:
        o = BUILD / (s + ".o")
        jobs.append(([cc, *COMMON, "-c", str(CSRC / s), "-o", str(o)], s))
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for w in ex.map(_compile, jobs):
            if verbose and w.strip():
                print(w)
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)],
                   check=True)
    os.replace(tmp, LIB)
    # the tests' command-line driver (the reference CLI's calls, linked to the library alone)
    cli = ROOT / "tests" / "cli"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Wno-format-truncation", "-o", str(cli / "dsmc_cli"), str(cli / "dsmc_cli.c"),
                    f"-L{PKG}", "-ldsmc_hip", "-Wl,-rpath,$ORIGIN/../../dsmc_suite_amd", "-lm"], check=True)
    if oracle:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        if Path("/root/reference/src/dsmc/ballistic.c").exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
        emu = ROOT / "tests" / "emu"
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o",
                        str(emu / "liblane_emu.so"), str(emu / "lane_emu.cpp")], check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
