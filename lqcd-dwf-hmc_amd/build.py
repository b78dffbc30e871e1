"""Builds libdwfhip.so (HIP, gfx950) in-tree, next to the Python binding.

    python lqcd-dwf-hmc_amd/build.py [-j N] [--verbose]

Every source is compiled with hipcc for gfx950 only; packed FP32 instructions
are disabled for the device code (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "dwf_amd", "libdwfhip.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["dwf_kernels.hip", "dwf_smear.hip", "dwf_force.hip", "dwf_lpc.hip", "dwf_batch.hip",
           "dwf_hmc.hip", "dwf_ptc.hip", "dwf_rng.cpp", "dwf_io.cpp", "dwf_api.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def compile_one(src, verbose=False):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return obj


def build(jobs=3, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_host_example(verbose)
    return OUT


EXAMPLE = os.path.join(HERE, "host", "example_solve")


def build_host_example(verbose=False):
    """The C++ host-side example driver (g++, links libdwfhip.so)."""
    libdir = os.path.dirname(OUT)
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "host"), os.path.join(HERE, "host", "example_solve.cpp"),
           "-L", libdir, "-ldwfhip", f"-Wl,-rpath,{libdir}", "-o", EXAMPLE + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host example build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(EXAMPLE + ".tmp", EXAMPLE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=3)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.verbose))


if __name__ == "__main__":
    main()
