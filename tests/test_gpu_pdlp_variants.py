"""Variants of the PDLP iteration in the single-LP engine:
  * restarts on PDLP's normalized duality gap (pm_pdhg_options.
    restart_metric = PM_RESTART_GAP): the restart rule compares the largest
    Lagrangian gap over the ball of radius ||z - z0|| around the candidate,
    divided by the radius, instead of the relative KKT error;
  * PDLP's adaptive step size (step_rule = PM_STEP_ADAPTIVE): plain PDHG
    whose step is accepted when it is at most ||dz||^2_omega / (2 |dy^T A
    dx|) and retried smaller otherwise, instead of the reflected Halpern
    iteration with the constant step 0.99 / ||A||_2.
Results are checked as for every PDLP solve: status, the oracle's FP64
relative KKT error of the returned (x, y), and the objective against the
generator's closed-form optimum.  The batch engine runs neither and
rejects both options."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gap_restarts_config0_to_1e8(capi, oracle, seed):
    lp = capi.generate(0, seed=seed)
    r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-8, max_iterations=100000,
                   check_frequency=64, restart_metric="gap")
    assert r.status == "OPTIMAL", r.status
    k = oracle.relative_kkt(lp, r.x, r.y)
    assert k["rel"] <= 1.2e-8, k
    assert abs(r.objective - lp.p_star) <= 1e-8 * (1 + abs(lp.p_star))
    assert r.restarts > 0


def test_gap_restarts_window_and_generated(capi, oracle):
    rh = capi.RollingHorizon()
    for lp in (rh.window(100), capi.generate(3, seed=12)):
        r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=100000,
                       check_frequency=64, restart_metric="gap")
        assert r.status == "OPTIMAL", r.status
        k = oracle.relative_kkt(lp, r.x, r.y)
        assert k["rel"] <= 1.2e-6, k


@pytest.mark.parametrize("seed", [1, 2])
def test_adaptive_step_config0_to_1e8(capi, oracle, seed):
    lp = capi.generate(0, seed=seed)
    r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-8, max_iterations=100000,
                   check_frequency=64, step_rule="adaptive")
    assert r.status == "OPTIMAL", r.status
    k = oracle.relative_kkt(lp, r.x, r.y)
    assert k["rel"] <= 1.2e-8, k
    assert abs(r.objective - lp.p_star) <= 1e-8 * (1 + abs(lp.p_star))


def test_adaptive_step_window(capi, oracle):
    lp = capi.RollingHorizon().window(0)
    r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=100000,
                   check_frequency=64, step_rule="adaptive")
    assert r.status == "OPTIMAL", r.status
    assert oracle.relative_kkt(lp, r.x, r.y)["rel"] <= 1.2e-6


def test_variants_rejected_by_batch_and_bad_values(capi):
    lp = capi.generate(0, seed=1)
    with pytest.raises(capi.PramaanError):
        capi.Batch([lp], tolerance=1e-6, restart_metric="gap")
    with pytest.raises(capi.PramaanError):
        capi.Batch([lp], tolerance=1e-6, step_rule="adaptive")
    with pytest.raises(capi.PramaanError):
        capi.solve(lp, algorithm="pdlp", termination="relative_kkt", restart_metric=7)
    with pytest.raises(capi.PramaanError):
        capi.solve(lp, algorithm="pdlp", termination="relative_kkt", step_rule=5)
