"""The kernels' per-node math headers (lbm_collide.h, lbm_particle_math.h)
built for the CPU with g++ under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/native/check_math.cpp): collision against
the FP64 oracle (1e-14; FP32 arithmetic 2e-6), moments, lattice tables, and
the interpolated bounce-back against the Lagrange interpolant of every
branch (tests/native/check_math_run.py)."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_kernel_math_under_sanitizers(tmp_path, oracle_lib):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    lib = str(tmp_path / "libcheck_math.so")
    r = subprocess.run([gxx, "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=undefined", "-o", lib,
                        os.path.join(HERE, "native", "check_math.cpp")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    asan = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(HERE, "native", "check_math_run.py"), lib],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "check_math: OK" in r.stdout, r.stdout + r.stderr
