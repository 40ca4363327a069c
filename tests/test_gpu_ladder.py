"""Precision ladder of this engine (pm_ladder_solve, built outside the
PRAMAAN tree, so with no CPU simplex linked): the reference's gate and
escalation reasons (src/gpu/precision_ladder.cpp:86-151) on rung 1, and the
FP64-refined GPU rung that answers where the reference would go to the CPU.
The reference's own ladder tests A-F run in test_gpu_reference_suite.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def test_loose_threshold_accepts_rung1(capi):
    # Wyndor (tests/test_gpu_pdhg.cpp:263-278): max 3x1 + 5x2, optimum 36
    lp = capi.LP.from_dense([[1, 0], [0, 2], [3, 2]], [3, 5], [-1e30] * 3, [4, 12, 18], [0, 0], [1e30, 1e30],
                            sense=1)
    r = capi.ladder_solve(lp, residual_threshold=1.0)
    assert r["stage"] == "GPU_FP32" and not r["escalated"] and r["escalation_reason"] == "none"
    assert r["gpu_status"] == "OPTIMAL" and r["final_status"] == "OPTIMAL"
    assert r["gpu_primal_residual"] <= 1.0 and r["gpu_dual_residual"] <= 1.0
    assert abs(r["final_objective"] - 36.0) <= 1e-2 and r["polish_unavailable"] == 0


def test_iteration_limit_escalates(capi):
    lp = capi.generate(0, seed=2)
    r = capi.ladder_solve(lp, residual_threshold=1.0, gpu_max_iterations=10)
    assert r["escalated"] and r["escalation_reason"] == "gpu_iteration_limit"
    assert r["gpu_iterations"] == 10
    # no CPU solver linked here: the best GPU point comes back, marked so
    assert r["polish_unavailable"] == 1 and r["stage"] == "GPU_FP32"


def test_refined_rung_answers_strict_threshold(capi):
    lp = capi.generate(0, seed=3)
    r = capi.ladder_solve(lp, residual_threshold=1e-8, refined_rung=True)
    assert r["escalated"] and r["escalation_reason"] in ("high_primal_residual", "high_dual_residual",
                                                         "gpu_iteration_limit")
    assert r["refined_rung_run"] == 1 and r["stage"] == "GPU_FP64_REFINED", r
    assert r["polish_unavailable"] == 0 and r["final_status"] == "OPTIMAL"
    assert r["final_primal_residual"] <= 1e-8 and r["final_dual_residual"] <= 1e-8
    assert abs(r["final_objective"] - lp.p_star) <= 1e-7 * (1 + abs(lp.p_star))


def test_refined_rung_on_afiro(capi):
    from pramaan_amd.mps import read_mps
    lp = read_mps(GOLDEN / "afiro.mps")
    r = capi.ladder_solve(lp, residual_threshold=1e-9, refined_rung=True)
    assert r["stage"] == "GPU_FP64_REFINED" and r["polish_unavailable"] == 0, r
    assert abs(r["final_objective"] + 464.75314285714285) <= 1e-6
