"""Builds libblaze2d_hip.so (gfx950) from blaze2d_amd/csrc with hipcc.

Run as ``python -m blaze2d_amd.build`` (or through ``__graft_entry__.build()``).
Objects go to blaze2d_amd/_build/, the shared library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libblaze2d_hip.so")
ARCH = os.environ.get("BLAZE2D_HIP_ARCH", "gfx950")

SOURCES = ["fft_ops.hip", "fft_big.hip", "blas.hip", "aux_kernels.hip", "dielectric.hip", "diagnostics.hip", "transition.hip", "engine.hip", "api.hip", "cnn.hip", "hostmath.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Packed FP32 instructions stay disabled in device code (-packed-fp32-ops).
# -fcx-limited-range: host std::complex products and quotients use the plain
# algebraic formulas instead of the NaN/Inf-recovering library calls.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fcx-limited-range", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-fcx-limited-range", "-x", "c++", "-c", os.path.join(CSRC, src),
               "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
