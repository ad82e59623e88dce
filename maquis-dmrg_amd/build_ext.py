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


# Host-side AddressSanitizer and UndefinedBehaviorSanitizer (the device code is built as usual): the
# CPU tests' instrumented build of the host planner, loaded with the clang ASan runtime preloaded
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
SANITIZED_BUILD = os.path.join(BUILD, "sanitized")


def _compile(src, build_dir=BUILD, flags=FLAGS):
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + flags + ["-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return obj


def build(verbose=False, sanitized=False):
    """Build libmqdmrg.so in-tree; sanitized=True builds the host-instrumented variant into
    _build/sanitized/libmqdmrg.so instead (test use only)."""
    build_dir = SANITIZED_BUILD if sanitized else BUILD
    flags = FLAGS + SANITIZE if sanitized else FLAGS
    lib = os.path.join(build_dir, "libmqdmrg.so") if sanitized else LIB
    os.makedirs(build_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, flags), SOURCES))
    LIB_OUT = lib
    tmp = LIB_OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fopenmp", "-o", tmp] + objs + [
        "-Wl,-rpath," + ROCM_LLVM_LIB, "-L" + ROCM_LLVM_LIB,
        "-Wl,-rpath," + ROCM_LIB, "-L" + ROCM_LIB, "-lrocblas", "-lhipblaslt"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB_OUT)
    if verbose:
        print("built", LIB_OUT)
    return LIB_OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
