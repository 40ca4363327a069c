"""Writes tests/golden/mrpl_rolling_reference.json: the per-day statuses and
objectives of the reference's own rolling-horizon MPS benchmark
(benchmark/rolling_horizon_mps_bench.cpp, compiled from the reference
sources into oracle/_ref by oracle/Makefile) on
tests/golden/mrpl_crude_blend.mps (the reference's benchmark/ model), 30
days, seed 42, perturbations 8 % / 6 %, tolerance 1e-8.  Cold days are the
reference RevisedSimplex, warm days its DualSimplex basis chain.  Timings
are dropped; only what the solvers computed is kept.

Usage: python tests/golden/make_mrpl_fixture.py
"""
import json
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
EXE = HERE.parent.parent / "oracle" / "_ref" / "rolling_horizon_mps_bench"


def main():
    out = subprocess.run([str(EXE), str(HERE / "mrpl_crude_blend.mps"), "--days", "30", "--seed", "42"],
                         capture_output=True, text=True, check=True).stdout
    run = json.loads(out)
    keep = ("day", "cold_objective", "warm_objective", "cold_status", "warm_status", "cold_iterations",
            "warm_iterations")
    fixture = {
        "source": "oracle/_ref/rolling_horizon_mps_bench mrpl_crude_blend.mps --days 30 --seed 42",
        "num_vars": run["num_vars"], "num_rows": run["num_rows"], "days": run["days"], "seed": run["seed"],
        "perturb_rhs": run["perturb_rhs"], "perturb_bounds": run["perturb_bounds"],
        "tolerance": run["tolerance"], "correctness": run["correctness"],
        "per_day": [{k: d[k] for k in keep} for d in run["per_day"]],
    }
    (HERE / "mrpl_rolling_reference.json").write_text(json.dumps(fixture, indent=1) + "\n")


if __name__ == "__main__":
    main()
