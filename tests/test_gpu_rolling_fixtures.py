"""Rolling-horizon windows against their FP64 optima: the six default-plan
and three perturbed-plan windows of tests/golden/fixtures.json (HiGHS
objectives written in the build container) solved to 1e-6 relative KKT in
one batch, from the cold start and from the trained network's warm start.
Every window must be OPTIMAL, its objective within 1e-6 relative of the
HiGHS optimum, and its (x, y) must have an FP64 relative KKT error of at
most 1.2e-6 when the oracle recomputes it from the returned vectors (the
engine's own check evaluates the same quantity on the device; 20 % head
room covers the FP64 summation order).  The same windows and tolerances
also go through the single-LP engine from the cold start."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FX = json.loads((Path(__file__).resolve().parent / "golden" / "fixtures.json").read_text())
OBJ_TOL = 1e-6
KKT_TOL = 1.2e-6


def fixture_windows(capi):
    rh = capi.RollingHorizon()
    rp = capi.RollingHorizon(seed=7, demand_noise=0.2, blend_noise=0.1, perturb_seed=11)
    out = [(f"default {w}", rh.window(int(w)), v) for w, v in sorted(FX["rolling"].items(), key=lambda t: int(t[0]))]
    out += [(f"perturbed {w}", rp.window(int(w)), v)
            for w, v in sorted(FX["rolling_perturbed"].items(), key=lambda t: int(t[0]))]
    return rh, out


def check(oracle, name, lp, r, want):
    assert r.status == "OPTIMAL", (name, r.status)
    assert abs(r.objective - want) <= OBJ_TOL * (1 + abs(want)), (name, r.objective, want)
    k = oracle.relative_kkt(lp, r.x, r.y)
    assert k["rel"] <= KKT_TOL, (name, k)


@pytest.mark.parametrize("start", ["cold", "learned"])
def test_fixture_windows_batch(capi, oracle, start):
    from pramaan_amd import rolling as R
    rh, wins = fixture_windows(capi)
    lps = [lp for _, lp, _ in wins]
    with capi.Batch(lps, tolerance=1e-6, max_iterations=200000, check_frequency=64) as b:
        if start == "learned":
            b.learned_warm_start(R.load(R.default_weights()), rh.cols_per_period, rh.rows_per_period)
        res, _, _ = b.run(want_vectors=True)
    for (name, lp, want), r in zip(wins, res):
        check(oracle, f"{start} {name}", lp, r, want)


def test_fixture_windows_single_lp_engine(capi, oracle):
    _, wins = fixture_windows(capi)
    for name, lp, want in wins[:3] + wins[-1:]:
        r = capi.solve(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=200000,
                       check_frequency=64)
        check(oracle, f"single {name}", lp, r, want)
        assert np.all(np.isfinite(r.x))
