"""The CPU oracle under sanitizers: oracle/sanitize_main.c drives every entry point of the
FP32 and FP64 restatements over scenes that reach their edge paths (colliding keys, one
table cell, one particle, a negative box minimum, the near-pressure term, walls, the density
map) and must exit cleanly under AddressSanitizer + UndefinedBehaviorSanitizer (serial,
every report fatal) and under ThreadSanitizer across the OpenMP loops (LLVM's OpenMP
runtime with its Archer tool).  Host code only; built by `make -C oracle sanitize`."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
BUILD = os.path.join(ORACLE, "build")
LLVM = "/opt/rocm/llvm"


@pytest.fixture(scope="module")
def sanitized():
    if shutil.which("gcc") is None or not os.path.exists(os.path.join(LLVM, "bin", "clang")):
        pytest.skip("gcc or the LLVM clang the ThreadSanitizer build uses is missing")
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return BUILD


def _run(exe, env):
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert out.strip().splitlines()[-1] == "sanitize: ok", out[-2000:]
    return out


def test_oracle_under_address_and_undefined_behaviour_sanitizers(sanitized):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    out = _run(os.path.join(sanitized, "sanitize_asan"), env)
    assert out.count("sanitize: ") == 6


def test_oracle_openmp_loops_under_thread_sanitizer(sanitized):
    archer = os.path.join(LLVM, "lib", "libarcher.so")
    if not os.path.exists(archer):
        pytest.skip("LLVM's Archer OpenMP tool is missing")
    env = dict(os.environ, OMP_NUM_THREADS="4", OMP_TOOL_LIBRARIES=archer,
               TSAN_OPTIONS="ignore_noninstrumented_modules=1 halt_on_error=1")
    out = _run(os.path.join(sanitized, "sanitize_tsan"), env)
    assert out.count("sanitize: ") == 6
