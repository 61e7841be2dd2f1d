"""The library's host modules (generators, species layout and coloring, BFM
files, R250) and the CPU oracle built with AddressSanitizer and
UndefinedBehaviorSanitizer and run over the reference's fixtures and
generated systems (tests/sanitize/).  CPU only; the GPU kernels are not part
of this build."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SAN = HERE / "sanitize"


def test_host_modules_and_oracle_under_asan_and_ubsan(tmp_path):
    if not shutil.which("g++") or not shutil.which("make"):
        pytest.skip("g++/make not available")
    r = subprocess.run(["make", "-s", "-j8", "-C", str(SAN)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([str(SAN / "_build" / "host_san"), str(HERE / "golden"), str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
