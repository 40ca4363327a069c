"""Rolling-horizon trace over a base MPS model, solved by GPU PDHG: the
protocol of the reference's benchmark/rolling_horizon_mps_bench.cpp with
the simplex replaced by this engine.

  1. the base model is read from MPS (pramaan_amd.mps.read_mps, the
     reference's parse_mps);
  2. day 0 is the base model, day d >= 1 perturbs every finite row bound
     and every positive finite variable upper bound by a factor drawn
     uniformly from [1 - p, 1 + p] (perturb_model, the reference's
     perturbModel, rolling_horizon_mps_bench.cpp:105-150: A and c never
     change; a perturbed bound that would cross its partner by less than
     1e-6 keeps its old value);
  3. cold: every day solved from scratch, one after another on the
     single-LP engine, and also all days at once as one batch;
  4. warm: day 0 cold, day d warm-started from day d-1's (x, y) (the
     chaining the reference does with simplex bases, :327-425);
  5. correctness: cold and warm objectives agree within tolerance x 1e4
     relative (the reference's gate, :432-445), and, when given, agree with
     the reference simplex's objectives for the same days.

The solves use PDLP with FP64 refinement to the relative KKT tolerance.
Output: one JSON object with the reference's aggregate and per-day keys
(times in ms; `basis_reused` reads "warm start kept" here), plus the
batched cold time.

Usage: python -m pramaan_amd.rolling_mps <model.mps> [--days N] [--seed S]
           [--perturb-rhs P] [--perturb-bounds P] [--tolerance T]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import _capi as P
from .mps import read_mps

INF = P.INF
MASK = (1 << 64) - 1
STATUS_CODE = {name: code for code, name in P.SOLVE_STATUS.items()}  # SolveStatus order, OPTIMAL = 0


class Lcg:
    """64-bit linear congruential generator of the reference's benchmarks
    (rolling_horizon_mps_bench.cpp:56-70): Knuth's MMIX multiplier and
    increment, uniform from the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next(self) -> int:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & MASK
        return self.state

    def uniform(self, lo: float, hi: float) -> float:
        t = float(self.next() >> 11) / float(1 << 53)
        return lo + t * (hi - lo)


def perturb_model(base: P.LP, day: int, seed: int = 42, perturb_rhs: float = 0.08,
                  perturb_bounds: float = 0.06) -> P.LP:
    """Day `day` of the trace (rolling_horizon_mps_bench.cpp:105-150)."""
    rng = Lcg(seed * 1000003 + day * 7919)
    rl, ru = base.row_lower.copy(), base.row_upper.copy()
    for r in range(base.m):
        up, lo = base.row_upper[r], base.row_lower[r]
        if -INF < up < INF:
            new_up = up * rng.uniform(1.0 - perturb_rhs, 1.0 + perturb_rhs)
            if new_up >= lo + 1e-6 or lo <= -INF:
                ru[r] = new_up
        if -INF < lo < INF:
            new_lo = lo * rng.uniform(1.0 - perturb_rhs, 1.0 + perturb_rhs)
            if new_lo <= ru[r] - 1e-6 or ru[r] >= INF:
                rl[r] = new_lo
    vu = base.var_upper.copy()
    for j in range(base.n):
        orig = base.var_upper[j]
        if 0.0 < orig < INF:
            new = orig * rng.uniform(1.0 - perturb_bounds, 1.0 + perturb_bounds)
            if new > base.var_lower[j] + 1e-6:
                vu[j] = new
    return dataclasses.replace(base, row_lower=rl, row_upper=ru, var_upper=vu, x_star=None, y_star=None,
                               p_star=None)


def day_models(base: P.LP, days: int, seed: int = 42, perturb_rhs: float = 0.08,
               perturb_bounds: float = 0.06) -> list[P.LP]:
    return [base] + [perturb_model(base, d, seed, perturb_rhs, perturb_bounds) for d in range(1, days)]


def primal_residual(lp: P.LP, x: np.ndarray) -> float:
    """Largest row or bound violation of x (the reference's
    computePrimalResidual, :155-178)."""
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_ptr))
    act = np.zeros(lp.m)
    np.add.at(act, rows, lp.values * x[lp.col_idx])
    v = [0.0]
    f = lp.row_lower > -INF
    v.append(float(np.max(lp.row_lower[f] - act[f], initial=0.0)))
    f = lp.row_upper < INF
    v.append(float(np.max(act[f] - lp.row_upper[f], initial=0.0)))
    f = lp.var_lower > -INF
    v.append(float(np.max(lp.var_lower[f] - x[f], initial=0.0)))
    f = lp.var_upper < INF
    v.append(float(np.max(x[f] - lp.var_upper[f], initial=0.0)))
    return max(v)


def run(mps_path, days=30, seed=42, perturb_rhs=0.08, perturb_bounds=0.06, tolerance=1e-8,
        max_iterations=200000, reference=None) -> dict:
    """The trace; `reference` optionally maps day -> the reference simplex's
    objective for that day."""
    base = read_mps(mps_path)
    models = day_models(base, days, seed, perturb_rhs, perturb_bounds)
    kw = dict(algorithm="pdlp", termination="relative_kkt", tolerance=tolerance, max_iterations=max_iterations,
              check_frequency=64)
    per = [dict(day=d, perturbation="base_case" if d == 0 else
                f"day{d}_rhs{perturb_rhs * 100:.0f}pct_bnd{perturb_bounds * 100:.0f}pct") for d in range(days)]
    ctx = P.Context(0)
    try:
        # cold, one day after another
        for d, lp in enumerate(models):
            ctx.load(lp, **kw)
            t0 = time.perf_counter()
            r = ctx.solve(want_vectors=True)
            per[d].update(cold_time_ms=(time.perf_counter() - t0) * 1e3, cold_objective=r.objective,
                          cold_iterations=r.iterations, cold_status=STATUS_CODE[r.status],
                          cold_primal_residual=primal_residual(lp, r.x), cold_rel_kkt=r.rel_kkt)
        # warm, chained
        x0 = y0 = None
        for d, lp in enumerate(models):
            k = dict(kw)
            if x0 is not None:
                k.update(x0=x0, y0=y0)
            ctx.load(lp, **k)
            t0 = time.perf_counter()
            r = ctx.solve(want_vectors=True)
            per[d].update(warm_time_ms=(time.perf_counter() - t0) * 1e3, warm_objective=r.objective,
                          warm_iterations=r.iterations, warm_status=STATUS_CODE[r.status],
                          warm_primal_residual=primal_residual(lp, r.x), warm_rel_kkt=r.rel_kkt,
                          basis_reused=d > 0, warm_fell_back=False)
            x0, y0 = r.x, r.y
    finally:
        ctx.close()
    # cold, all days in one batch
    with P.Batch(models, tolerance=tolerance, max_iterations=max_iterations, check_frequency=64) as b:
        b.run()
        t0 = time.perf_counter()
        bres, _, _ = b.run()
        batch_ms = (time.perf_counter() - t0) * 1e3
    all_correct = True
    for d, p in enumerate(per):
        ref = max(abs(p["cold_objective"]), 1.0)
        p["objective_abs_diff"] = abs(p["cold_objective"] - p["warm_objective"])
        p["correctness_ok"] = p["objective_abs_diff"] / ref <= tolerance * 1e4
        p["batch_objective"] = bres[d].objective
        p["batch_status"] = STATUS_CODE[bres[d].status]
        if reference is not None and d in reference:
            rv = reference[d]
            p["reference_objective"] = rv
            p["reference_rel_diff"] = max(abs(p["cold_objective"] - rv), abs(p["warm_objective"] - rv),
                                          abs(p["batch_objective"] - rv)) / max(abs(rv), 1.0)
            p["correctness_ok"] = p["correctness_ok"] and p["reference_rel_diff"] <= tolerance * 1e4
        all_correct = all_correct and p["correctness_ok"]
    cold_ms = sum(p["cold_time_ms"] for p in per)
    warm_ms = sum(p["warm_time_ms"] for p in per)
    cold_it = sum(p["cold_iterations"] for p in per)
    warm_it = sum(p["warm_iterations"] for p in per)
    cold_opt = sum(p["cold_status"] == 0 for p in per)
    warm_opt = sum(p["warm_status"] == 0 for p in per)
    return {
        "instance": str(mps_path), "num_vars": base.n, "num_rows": base.m, "days": days, "seed": seed,
        "perturb_rhs": perturb_rhs, "perturb_bounds": perturb_bounds, "tolerance": tolerance,
        "solver": "GPU PDLP (FP32 iterations, FP64 refinement and relative KKT)",
        "cold_cumulative_ms": cold_ms, "warm_cumulative_ms": warm_ms,
        "cumulative_ratio": cold_ms / warm_ms if warm_ms > 0 else 0.0,
        "cold_batched_ms": batch_ms,
        "cold_total_iters": cold_it, "warm_total_iters": warm_it,
        "iteration_ratio": cold_it / warm_it if warm_it > 0 else 0.0,
        "cold_optimal_days": cold_opt, "warm_optimal_days": warm_opt,
        "batch_optimal_days": sum(p["batch_status"] == 0 for p in per),
        "warm_start_attempts": days - 1, "warm_start_successes": sum(p["basis_reused"] for p in per),
        "warm_start_fallbacks": 0,
        "benchmark_valid": all_correct and cold_opt == days and warm_opt == days,
        "correctness": all_correct,
        "per_day": per,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("mps")
    ap.add_argument("--days", type=int, default=30)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--perturb-rhs", type=float, default=0.08)
    ap.add_argument("--perturb-bounds", type=float, default=0.06)
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--reference", help="JSON of the reference run: per_day[d].cold_objective is compared")
    a = ap.parse_args(argv)
    if a.days < 1:
        print("--days must be >= 1", file=sys.stderr)
        return 1
    ref = None
    if a.reference:
        with open(a.reference) as f:
            ref = {int(p["day"]): float(p["cold_objective"]) for p in json.load(f)["per_day"]}
    out = run(a.mps, a.days, a.seed, a.perturb_rhs, a.perturb_bounds, a.tolerance, reference=ref)
    print(json.dumps(out, indent=2))
    return 0 if out["benchmark_valid"] else 1


if __name__ == "__main__":
    sys.exit(main())
