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

Callers of the drop-in built the same way:
  * the reference CLI (src/cli/solve_main.cpp, its GPU branch) with the
    reference's tests/test_cli_precision_ladder.cpp, which runs
    `pramaan-solve model.mps "" --gpu` and requires the GPU path;
  * the reference's benchmark/precision_ladder_bench.cpp on afiro (its
    direct GPU row and the threshold sweep through the reference
    controller);
  * this project's tests/cxx/test_backend_hip.cpp: make_hip_backend()
    through the reference's Backend header, four threads on one instance.

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


def run_ref(name, tmp_path, *args, also=()):
    exe = REF / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (needs the reference sources at build time)")
    (tmp_path / "tests").mkdir()
    os.symlink(ROOT / "tests" / "golden", tmp_path / "tests" / "data")
    for other in also:  # programs the test runs from its working directory
        os.symlink(REF / other, tmp_path / other)
    p = subprocess.run([str(exe), *args], cwd=tmp_path, capture_output=True, text=True, timeout=900)
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


def test_reference_cli_gpu_path(tmp_path):
    p = run_ref("test_cli_precision_ladder", tmp_path, also=("pramaan-solve",))
    assert p.returncode == 0 and "CLI routed to GPU" in p.stdout


def test_reference_cli_solves_afiro_on_gpu(tmp_path):
    p = run_ref("pramaan-solve", tmp_path, "tests/data/afiro.mps", "", "--gpu")
    assert p.returncode == 0, p.stderr
    assert "GPU PDHG candidate" in p.stdout and "Status: OPTIMAL" in p.stdout
    assert "Falling back to CPU" not in p.stderr


def test_reference_precision_ladder_bench(tmp_path):
    p = run_ref("precision_ladder_bench", tmp_path, "afiro")
    assert p.returncode == 0
    rows = [r.split(",") for r in p.stdout.splitlines() if r.startswith("afiro,")]
    assert any(r[2] == "GPU_FP32" and r[-1] == "1" for r in rows), p.stdout


def test_hip_backend_concurrent_calls(tmp_path):
    p = run_ref("test_backend_hip", tmp_path)
    assert p.returncode == 0 and "ALL PASSED" in p.stdout
