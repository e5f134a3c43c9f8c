"""Builds libsph_hip.so (gfx950) in-tree with hipcc.

Usage: python particle-fluid-simulation_amd/build.py [--jobs N]

Every HIP translation unit is compiled with
  -ffp-contract=off         no FMA contraction: each float operation rounds as written,
                            matching the reference's operation order bit for bit
  -packed-fp32-ops disabled the compiler may not form v_pk_*_f32 instructions
and IEEE division / sqrt (hipcc's default), so results do not depend on code generation.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIBNAME = "libsph_hip.so"
ARCH = "gfx950"

SOURCES = ["sph_kernels.hip", "sph_tiles.hip", "sph_brick.hip", "radix_sort.hip", "sph_layout.hip", "sph_density_map.hip",
           "sph_world.hip", "sph_batch.hip", "sph_server.hip", "sph_host.hip", "sph_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-result",
]


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = lib_path() + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib_path())
    if verbose:
        print(f"built {lib_path()}")
    build_driver(verbose)
    return lib_path()


def build_driver(verbose: bool = False) -> str:
    """The headless C driver (driver/sph_run.c) linked against libsph_hip.so."""
    out = os.path.join(HERE, "bin", "sph_run")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = ["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-Wall", f"-I{inc}", os.path.join(HERE, "driver", "sph_run.c"),
           f"-L{LIBDIR}", "-lsph_hip", f"-Wl,-rpath,$ORIGIN/../lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {out}")
    return out


def clean() -> None:
    shutil.rmtree(OBJDIR, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
