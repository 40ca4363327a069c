"""The MPS rolling-horizon trace on the GPU (pramaan_amd/rolling_mps.py, the
protocol of the reference's benchmark/rolling_horizon_mps_bench.cpp): 30
perturbed days of mrpl_crude_blend.mps solved cold one by one, cold as one
batch, and warm-chained, each to 1e-8 relative KKT.  Every day is optimal on
every path and every objective agrees with the reference's own simplex
(tests/golden/mrpl_rolling_reference.json) to 1e-6 relative; the
reference's gate between cold and warm (tolerance x 1e4) holds.
"""
import json
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_rolling_mps_trace_matches_reference():
    from pramaan_amd import rolling_mps as R
    fx = json.loads((GOLDEN / "mrpl_rolling_reference.json").read_text())
    ref = {p["day"]: p["cold_objective"] for p in fx["per_day"]}
    out = R.run(GOLDEN / "mrpl_crude_blend.mps", days=30, seed=42, tolerance=1e-8, reference=ref)
    print({k: v for k, v in out.items() if k != "per_day"})
    assert out["benchmark_valid"] and out["correctness"]
    assert out["cold_optimal_days"] == out["warm_optimal_days"] == out["batch_optimal_days"] == 30
    for p in out["per_day"]:
        assert p["reference_rel_diff"] <= 1e-6, p
        assert p["cold_primal_residual"] <= 1e-6 * (1 + abs(p["cold_objective"])), p
