"""The CPU-only build of the drop-in: the reference's
tests/test_precision_ladder.cpp compiled without PRAMAAN_ENABLE_CUDA over this
project's controller compiled without PRAMAAN_ENABLE_HIP (oracle/Makefile
target test_precision_ladder_cpu).  It links no GPU code, and its results
are the reference's kGpuDisabled ones (src/gpu/precision_ladder.cpp:156-173,
expected by the test's #else branches)."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "oracle" / "_ref" / "test_precision_ladder_cpu"


def test_cpu_only_ladder_build(tmp_path):
    if not EXE.exists():
        pytest.skip(f"{EXE} not built (needs the reference sources at build time)")
    libs = subprocess.run(["ldd", str(EXE)], capture_output=True, text=True).stdout
    assert "pramaan" not in libs and "amdhip" not in libs, libs
    (tmp_path / "tests").mkdir()
    os.symlink(ROOT / "tests" / "golden", tmp_path / "tests" / "data")
    p = subprocess.run([str(EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0 and "TEST SUITE PASSED" in p.stdout, p.stdout[-3000:]
