"""libteuk's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(no GPU): tests/sanitize/host_checks.cpp drives the grid inversions, the
coefficient, projection and phase tables, the per-step source preparation
and T4 coefficients, the trajectory generator over steep, fast, near-polar
and high-spin worldlines, and the stage kernels' theta-chunk layout and
block order (every row spanned and computed once, every tile visited once),
built with g++ -fsanitize=address,undefined
from the same sources as the library.  Any out-of-bounds access, use after
free, signed overflow or other undefined operation aborts the run, and every
output must be finite.  (The reference has no such checks; its host code
writes qq[M] out of bounds, proj.c:45, which the restatement avoids.)"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bh_ringdown_amd", "csrc")
HIP_INCLUDE = "/opt/rocm/include"


def test_host_code_clean_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(os.path.join(HIP_INCLUDE, "hip")):
        pytest.skip("needs g++ and the HIP headers")
    exe = tmp_path / "host_checks"
    build = subprocess.run(
        [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-D__HIP_PLATFORM_AMD__", "-I", HIP_INCLUDE, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
         "-o", str(exe), os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"),
         os.path.join(CSRC, "teuk_host.cpp"), os.path.join(CSRC, "teuk_traj.cpp")],
        capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-3000:]
    # verify_asan_link_order=0: the environment may preload libraries ahead of the sanitizer runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout + run.stderr)[-3000:]
    assert "host checks: 0 failures" in run.stdout
