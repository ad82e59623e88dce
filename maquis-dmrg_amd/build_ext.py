"""Build the gfx950 engine library ``libmqdmrg.so`` in-tree with hipcc.

All sources (host planner in C++, kernels and C-ABI in HIP) go into one shared library next to
this file.  HIP code is compiled for gfx950 only, with packed FP32 instructions disabled.  The
plain dense GEMMs of the grouped S stage call rocBLAS (linked from /opt/rocm/lib).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmqdmrg.so")

SOURCES = [
    "csrc/host/schedule.cpp",
    "csrc/device/kernels.hip",
    "csrc/device/gemm.hip",
    "csrc/device/s_stage.hip",
    "csrc/device/t_stage.hip",
    "csrc/device/f_stage.hip",
    "csrc/device/engine.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LLVM_LIB = "/opt/rocm/llvm/lib"
ROCM_LIB = "/opt/rocm/lib"
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fopenmp",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp] + objs + [
        "-Wl,-rpath," + ROCM_LLVM_LIB, "-L" + ROCM_LLVM_LIB,
        "-Wl,-rpath," + ROCM_LIB, "-L" + ROCM_LIB, "-lrocblas", "-lhipblaslt"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
