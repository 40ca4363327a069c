"""Restarts on PDLP's normalized duality gap (pm_pdhg_options.restart_metric
= PM_RESTART_GAP, single-LP engine): the restart rule compares the largest
Lagrangian gap over the ball of radius ||z - z0|| around the candidate,
divided by the radius, instead of the relative KKT error.  It only changes
when restarts happen, so results are checked as for every PDLP solve:
status, the oracle's FP64 relative KKT error of the returned (x, y), and
the objective against the generator's closed-form optimum.  The batch
engine restarts on the KKT error and rejects the option."""
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


def test_gap_restarts_rejected_by_batch_and_bad_values(capi):
    lp = capi.generate(0, seed=1)
    with pytest.raises(capi.PramaanError):
        capi.Batch([lp], tolerance=1e-6, restart_metric="gap")
    with pytest.raises(capi.PramaanError):
        capi.solve(lp, algorithm="pdlp", termination="relative_kkt", restart_metric=7)
