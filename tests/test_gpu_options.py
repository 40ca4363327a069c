"""Every algorithm / termination pair on the GPU: the accepted ones solve
to their own test, the rejected ones fail with PM_ERR_INVALID_ARG.  Under
PM_TERM_REFERENCE the PDLP solve must meet the reference's pv and dv
(pdhg_solver.cu:124-157,356) as well as the relative KKT tolerance; the
same solve under PM_TERM_RELATIVE_KKT runs beside it and can stop no
later."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_pdlp_reference_termination_meets_pv_dv(capi, oracle):
    lp = capi.generate(0, seed=2)
    tol = 1e-6
    kkt = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=tol, check_frequency=64)
    ref = capi.solve(lp, algorithm="pdlp", termination="reference", tolerance=tol, check_frequency=64)
    assert kkt.status == ref.status == "OPTIMAL"
    assert kkt.rel_kkt <= tol and ref.rel_kkt <= tol
    assert ref.primal_residual <= tol and ref.dual_residual <= tol, (ref.primal_residual, ref.dual_residual)
    assert ref.iterations >= kkt.iterations
    assert abs(ref.objective - lp.p_star) <= 1e-6 * (1 + abs(lp.p_star))


def test_reference_pair_and_rejections(capi):
    lp = capi.generate(0, seed=1)
    r = capi.solve(lp, algorithm="reference", termination="reference", tolerance=1e-4, max_iterations=20000)
    assert r.status == "OPTIMAL" and r.primal_residual <= 1e-4 and r.dual_residual <= 1e-4
    for kw in (dict(termination="relative_kkt"), dict(x0=np.zeros(lp.n))):
        with pytest.raises(capi.PramaanError) as e:
            capi.solve(lp, algorithm="reference", **kw)
        assert e.value.code == 1


def test_pdlp_warm_start_accepted(capi):
    lp = capi.generate(0, seed=3)
    cold = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-7, check_frequency=64)
    warm = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-7, check_frequency=64,
                      x0=cold.x, y0=cold.y)
    assert cold.status == warm.status == "OPTIMAL" and warm.iterations <= cold.iterations
