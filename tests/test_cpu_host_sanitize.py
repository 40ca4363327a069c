"""This project's host-side C/C++ code (synthetic and rolling-horizon
generators, the network's feature builder, certificates and ModelIR) built
with AddressSanitizer and UndefinedBehaviorSanitizer (tests/cxx/Makefile,
target `sanitize`; g++ only, no GPU code) and driven through its C-ABI by
tests/cxx/test_host_sanitize.cpp, which checks results as well as running
clean: any sanitizer report aborts the run."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cxx" / "_build" / "test_host_sanitize"


def test_host_code_under_asan_and_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    b = subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cxx"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([str(EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0 and "ALL PASSED" in p.stdout, p.stdout[-4000:] + p.stderr[-4000:]
