"""The perturbed daily models of the MPS rolling-horizon trace
(pramaan_amd/rolling_mps.py, the reference's perturbModel,
benchmark/rolling_horizon_mps_bench.cpp:105-150), on the CPU.

Pinned against the reference's own benchmark, compiled from its sources:
tests/golden/mrpl_rolling_reference.json holds its per-day simplex
objectives on mrpl_crude_blend.mps (30 days, seed 42).  HiGHS (scipy,
test-only) solves this module's day models; the objectives must agree to
1e-8 relative, which a different seed, draw order or acceptance rule would
break wherever the bound it moves is active.  The rules are checked
directly too: A and c never change, absent bounds stay absent, each moved
bound moves by a factor within [1 - p, 1 + p], day 0 is the base model.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
INF = 1e30


@pytest.fixture(scope="module")
def trace():
    from pramaan_amd import rolling_mps as R
    from pramaan_amd.mps import read_mps
    base = read_mps(GOLDEN / "mrpl_crude_blend.mps")
    return base, R.day_models(base, 30, seed=42)


def test_days_match_reference_objectives(trace):
    pytest.importorskip("scipy")
    sys.path.insert(0, str(GOLDEN))
    from make_fixtures import highs
    fx = json.loads((GOLDEN / "mrpl_rolling_reference.json").read_text())
    base, models = trace
    assert (base.m, base.n) == (fx["num_rows"], fx["num_vars"])
    for d, lp in enumerate(models):
        ref = fx["per_day"][d]
        assert ref["cold_status"] == 0
        got = highs(lp)
        assert abs(got - ref["cold_objective"]) <= 1e-8 * max(1.0, abs(ref["cold_objective"])), d


def test_perturbation_rules(trace):
    base, models = trace
    m0 = models[0]
    assert np.array_equal(m0.row_lower, base.row_lower) and np.array_equal(m0.var_upper, base.var_upper)
    moved = 0
    for lp in models[1:]:
        assert np.array_equal(lp.values, base.values) and np.array_equal(lp.c, base.c)
        assert np.array_equal(lp.var_lower, base.var_lower)
        for new, old, p in ((lp.row_lower, base.row_lower, 0.08), (lp.row_upper, base.row_upper, 0.08),
                            (lp.var_upper, base.var_upper, 0.06)):
            absent = np.abs(old) >= INF
            assert np.array_equal(new[absent], old[absent])
            f = (~absent) & (old != 0.0) & (new != old)
            moved += int(f.sum())
            ratio = new[f] / old[f]
            assert np.all(ratio >= 1.0 - p - 1e-12) and np.all(ratio <= 1.0 + p + 1e-12)
        assert np.all(lp.row_lower <= lp.row_upper)
    assert moved > 0
