"""The reference's own GPU test programs, built from the reference sources by
oracle/Makefile (target `reftests`) against the reference headers, with this
engine's drop-in gpu::PdhgSolver and test launchers in place of the CUDA
target.  They print "TEST SUITE PASSED" and exit 0 on success
(tests/test_gpu_pdhg.cpp:524-545, tests/test_pdhg_kernel.cpp:243-259,
tests/test_precision_ladder.cpp:497-520, tests/test_presolve_scaling.cpp:
670-690).  The precision-ladder tests A-F run twice: over the reference's
own controller (an unchanged caller of the drop-in solver) and over this
project's controller.  The presolve/scaling suite runs the reference's
fixed-variable presolve and simplex with this project's ruiz_scale /
unscale_solution (the passes on the device).

The programs read tests/data/{afiro,adlittle}.mps relative to their working
directory; the run directory links that path to tests/golden.
"""
import os
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"


def run_ref(name, tmp_path):
    exe = REF / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (needs the reference sources at build time)")
    (tmp_path / "tests").mkdir()
    os.symlink(ROOT / "tests" / "golden", tmp_path / "tests" / "data")
    p = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=900)
    print(p.stdout[-4000:])
    print(p.stderr[-4000:])
    return p


def test_reference_test_pdhg_kernel(tmp_path):
    p = run_ref("test_pdhg_kernel", tmp_path)
    assert p.returncode == 0 and "TEST SUITE PASSED" in p.stdout


def test_reference_test_gpu_pdhg(tmp_path):
    p = run_ref("test_gpu_pdhg", tmp_path)
    assert p.returncode == 0 and "TEST SUITE PASSED" in p.stdout


def test_reference_ladder_tests_reference_controller(tmp_path):
    p = run_ref("test_precision_ladder_ref", tmp_path)
    assert p.returncode == 0 and "TEST SUITE PASSED" in p.stdout


def test_reference_ladder_tests_engine_controller(tmp_path):
    p = run_ref("test_precision_ladder", tmp_path)
    assert p.returncode == 0 and "TEST SUITE PASSED" in p.stdout


def test_reference_presolve_scaling_tests(tmp_path):
    p = run_ref("test_presolve_scaling", tmp_path)
    assert p.returncode == 0 and "TEST SUITE PASSED" in p.stdout
