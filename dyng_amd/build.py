"""Build libdyng_hip.so (the HIP sssp path, gfx950) in-tree with hipcc.

Usage: python -m dyng_amd.build [--jobs N]. The library lands next to this file; the objects go
to dyng_amd/_build/. Every HIP translation unit is compiled with packed-FP32 instructions off.
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
OUT = os.path.join(HERE, "libdyng_hip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["graph.hip", "sssp.hip", "scheduler.hip", "mosp.hip", "api.cpp", "hyper.cpp"]
ARCH = os.environ.get("DYNG_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP sssp path)")


def flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(HERE, "..", "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + flags()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
    return obj


GEN_OUT = os.path.join(HERE, "libdyng_gen.so")


def build_generators() -> str:
    """The host-side input generators (C++/OpenMP, g++)."""
    tmp = GEN_OUT + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared",
           "-std=c++17", os.path.join(CSRC, "generators.cpp"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building libdyng_gen.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, GEN_OUT)
    return GEN_OUT


COPY_OUT = os.path.join(HERE, "hbm_copy")


def build_copy_probe() -> str:
    """hbm_copy: the known-size copy that calibrates the HBM traffic counters."""
    tmp = COPY_OUT + ".tmp"
    cmd = [hipcc(), "-x", "hip", os.path.join(CSRC, "hbm_copy.hip"), "-o", tmp] + flags()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building hbm_copy failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, COPY_OUT)
    return COPY_OUT


def build(jobs: int = 4) -> str:
    build_generators()
    build_copy_probe()
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libdyng_hip.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args()
    print(build(args.jobs))
    return 0


if __name__ == "__main__":
    sys.exit(main())
