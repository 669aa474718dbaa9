"""The library's host-side C++ under AddressSanitizer and
UndefinedBehaviorSanitizer, without a GPU.

`make sanitize` compiles the host sources (table builders, dense and
split-fp16 operator packing, argument checks, exact-file IO, the C ABI) with
both sanitizers on the host pass and links them with the device objects into
host_check (laplace-fmm_amd/csrc/host_check.cpp), which checks the tables
(Wigner orthogonality, finiteness, prefix bounds), the exact-file round trip
and the null-context error path of every entry point.  Any sanitizer report
aborts the program (-fno-sanitize-recover=all).
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "laplace-fmm_amd", "csrc")
EXE = os.path.join(ROOT, "laplace-fmm_amd", "_build_san", "host_check")


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_host_code_under_asan_and_ubsan(tmp_path):
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-s", "-j" + jobs, "sanitize"], cwd=CSRC, capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, (out.stdout[-4000:], out.stderr[-4000:])
    assert "host_check ok" in out.stdout
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr
