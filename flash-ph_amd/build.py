"""Build libflashph.so (gfx950) in-tree from flash-ph_amd/csrc, and the host
pipeline libflashph_cpu.so from flash-ph_amd/cpu.

Each .hip file is compiled to an object with hipcc, in parallel, then linked
into flash-ph_amd/flash_ph/_lib/libflashph.so; the CPU library is one g++
invocation (OpenMP).  Always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "flash_ph", "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libflashph.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions: every FP32 op is a plain, separately rounded one
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


CPU_SRC = os.path.join(HERE, "cpu", "fph_cpu.cpp")
CPU_LIB = os.path.join(OUT_DIR, "libflashph_cpu.so")
CXX = os.environ.get("CXX", "g++")
# -ffp-contract=off: the only fused operations are the explicit fmaf calls
# of the distance chain (hardware FMA with -mfma, single rounding either way)
CPU_FLAGS = ["-O3", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-mfma", "-mavx2",
             "-fPIC", "-shared", "-Wall", "-Wno-unused-function"]


def build_cpu() -> str:
    """libflashph_cpu.so: the host pipeline for clouds in host memory."""
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = CPU_LIB + ".tmp"
    r = subprocess.run([CXX, *CPU_FLAGS, "-o", tmp, CPU_SRC], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {CPU_SRC}:\n{r.stderr[-6000:]}")
    os.replace(tmp, CPU_LIB)
    return CPU_LIB


def build(jobs: int | None = None) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4) // 2))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    build_cpu()
    return LIB


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else None))
