"""The host planner (csrc/host/schedule.cpp and the engine's host code, through libmqdmrg.so's
C-ABI) and the oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (host code only;
build_ext.build(sanitized=True), `make -C oracle sanitized`), running the planner-vs-oracle and
descriptor-rejection tests in a subprocess that preloads clang's ASan runtime.  Any out-of-bounds
access, use after free or undefined behaviour in the planner, the boundary-index builders or the
oracle stops that run with a report."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "maquis-dmrg_amd")
sys.path.insert(0, PKG)


def _runtime():
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def _stale(target, sources):
    return not os.path.exists(target) or any(os.path.getmtime(s) > os.path.getmtime(target) for s in sources)


@pytest.fixture(scope="module")
def sanitized_libs():
    rt = _runtime()
    if rt is None:
        pytest.skip("clang ASan runtime not found under /opt/rocm/llvm")
    import build_ext
    srcs = [os.path.join(PKG, s) for s in build_ext.SOURCES] + glob.glob(os.path.join(PKG, "csrc", "*", "*.hpp")) + \
        [os.path.join(ROOT, "include", "mq_sigma.h")]
    lib = os.path.join(build_ext.SANITIZED_BUILD, "libmqdmrg.so")
    if _stale(lib, srcs):
        build_ext.build(sanitized=True)
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitized"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return rt, lib, os.path.join(ROOT, "oracle", "_build", "sanitized", "libmqoracle.so")


def test_planner_and_oracle_clean_under_asan_and_ubsan(sanitized_libs):
    rt, lib, olib = sanitized_libs
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": rt,
        # leaks and allocator-pairing reports of the interpreter and its extension modules are not ours
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:alloc_dealloc_mismatch=0:detect_odr_violation=0:"
                        "verify_asan_link_order=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "MQ_LIB_PATH": lib,
        "MQ_ORACLE_LIB": olib,
    })
    sel = "schedule or step_indices or index_only or rejects or oracle_sigma or boundary_steps or hermitian or exports"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_cpu_oracle_and_planner.py"), "-k", sel],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, out
    assert "passed" in r.stdout and "AddressSanitizer" not in out and "runtime error" not in out, out
