#!/usr/bin/env python3
"""Benchmark: seconds to 1e-6 relative KKT on the 500-window rolling-horizon
blending sequence (24-period windows shifted by one period; 5,160 rows x
9,600 columns and 109k nonzeros per window), learned warm start vs plain
cold-start PDHG, on one GPU, plus the SpMV HBM bandwidth.

A step solves all 500 windows from scratch with every window already
resident in HBM (one pm_batch):
  learned = the network's graph and features built on the device from the
            resident windows, one forward pass for all windows, each
            prediction judged against the cold start by its FP64 relative
            KKT error, then one batched, graph-captured PDHG solve
            (reflected Halpern, FP32 iterations, FP64 checks and
            refinement) of all windows to 1e-6;
  cold    = the same solve from x = 0, y = 0.
`value` is the learned step's mean wall time (features + inference + solve);
the cold step is timed the same way beside it.  The network weights are
the ones trained by `python -m pramaan_amd.rolling` (pramaan_amd/data/).

roofline: the dominant iterate kernel of the batched solve (A^T y with the
fused primal update, or A xe with the fused dual update, over all 500
windows), timed with HIP events on the batch's stream; algorithmic bytes
per launch = sum over windows of 8 nnz + 4 (n+1) + 4 m + 28 n (A^T) or
8 nnz + 4 (m+1) + 4 n + 20 m (A).  spmv_lp1: the same for the single-LP
engine's fused SpMV on the 500k x 1M LP (12M nonzeros).
plain_pdhg_gpu: plain cold-start PDHG on the GPU, i.e. the reference
algorithm restated on the device (PM_ALG_REFERENCE), run to its own end on
a few sample windows one after another; the FP64 relative KKT error of
each final point comes from the batch engine's warm-start judge.
cpu_baseline: the oracle's restatement of the reference PDHG (FP32, its
own termination test pv, dv <= tol and its default 100,000-iteration
budget) run to its end on a sample of windows spread over the sequence,
one window per host thread, all threads at once, timed by the wall clock;
value = the sample's wall seconds x (windows / sample size), i.e. the
whole sequence on the same threads.  How many windows reached the
tolerance and their iteration counts are reported beside it, and a
single-thread run of one window.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload lp1        (the single 500k x 1M LP instead; --lp tenth for
                                              50k x 100k, --dense-rows K appends K rows over
                                              every column)
       python bench.py --workload batch4096  (4,096 generated LPs in one batch)
       python bench.py --workload ruiz       (Ruiz scaling of the 500k x 1M LP)
       python bench.py --workload train      (network training on 20,000 windows)
       python bench.py --workload traffic    (dispatches for rocprofv3 counters)
       python bench.py --workload starts     (solve time by warm start: learned, exact, noisy, active set)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
WEIGHTS = ROOT / "pramaan_amd" / "data" / "warmstart_gnn.npz"
METRIC = "s to 1e-6 rel KKT, learned warm-start vs plain PDHG on 1 GPU; SpMV HBM GB/s"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel: str):
    """HBM bytes per launch of an iterate kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    path = ROOT / "profiles" / "batch_iterate_rolling500_traffic.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def oracle_seconds_per_iteration(lp, budget_s: float):
    """Restated reference PDHG (oracle, FP32, 1 thread): seconds per
    iteration on lp, setup subtracted, over a bounded run."""
    from oracle import binding as O
    t0 = time.perf_counter()
    O.pdhg_solve(lp, tol=0.0, max_iterations=0, check_frequency=10**9)
    setup = time.perf_counter() - t0
    iters = 20
    t0 = time.perf_counter()
    O.pdhg_solve(lp, tol=0.0, max_iterations=iters, check_frequency=10**9)
    per_it = max((time.perf_counter() - t0 - setup) / iters, 1e-9)
    iters = max(20, min(200000, int(budget_s / per_it)))
    t0 = time.perf_counter()
    O.pdhg_solve(lp, tol=0.0, max_iterations=iters, check_frequency=10**9)
    return max((time.perf_counter() - t0 - setup) / iters, 1e-9), iters


def reference_residuals(lp, x, y, tol):
    """The reference's convergence measures of (x, y) (ComputeFp64Residuals,
    src/gpu/pdhg_solver.cu:107-161): pv = largest row or bound violation of
    x in FP64; dv = largest reduced-cost violation c + A^T y (c and A rounded
    to FP32 as the reference holds them, minimisation form) allowed by where
    x sits within tol of its bounds, divided by max(1, |c|inf)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_ptr))
    act = np.zeros(lp.m)
    np.add.at(act, rows, lp.values * x[lp.col_idx])
    pv = max(0.0, float(np.max(lp.row_lower - act, initial=0.0)), float(np.max(act - lp.row_upper, initial=0.0)),
             float(np.max(lp.var_lower - x, initial=0.0)), float(np.max(x - lp.var_upper, initial=0.0)))
    c = ((-1.0 if lp.sense == 1 else 1.0) * lp.c).astype(np.float32).astype(np.float64)
    aty = np.zeros(lp.n)
    np.add.at(aty, lp.col_idx, lp.values.astype(np.float32).astype(np.float64) * y[rows])
    rc = c + aty
    at_lb = x <= lp.var_lower + tol
    at_ub = x >= lp.var_upper - tol
    viol = np.where(~at_lb & ~at_ub, np.abs(rc),
                    np.where(at_lb & ~at_ub, np.maximum(0.0, -rc),
                             np.where(at_ub & ~at_lb, np.maximum(0.0, rc), 0.0)))
    dv = float(np.max(viol, initial=0.0)) / max(1.0, float(np.max(np.abs(c), initial=0.0)))
    return pv, dv


def cpu_threads() -> int:
    """Host threads for the parallel CPU baseline: OMP_NUM_THREADS when set
    (the box's CPU share), at most the visible cores."""
    n = os.cpu_count() or 1
    try:
        return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))
    except ValueError:
        return n


def oracle_to_termination(lps, threads: int, tol: float, max_iterations: int = 100000):
    """The oracle's restated reference PDHG (pdhg_solver.cu:169-411: FP32
    iterations, FP64 pv / dv check every 100 iterations, its own
    termination test and restarts) run to its end on every LP, `threads`
    LPs at a time, one per host thread.  Returns (wall seconds, results);
    each result also carries the CPU seconds of its own thread."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import binding as O

    def one(lp):
        t = time.thread_time()
        r = O.pdhg_solve(lp, tol=tol, max_iterations=max_iterations)[0]
        r["thread_seconds"] = time.thread_time() - t
        return r
    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        res = list(ex.map(one, lps))
        wall = time.perf_counter() - t0
    return wall, res


def spmv_lp1(P, reps: int) -> dict:
    lp = P.generate(1, seed=1)
    with P.Context(0) as ctx:
        ctx.load(lp, algorithm="pdlp", termination="relative_kkt", tolerance=1e-6, max_iterations=64,
                 check_frequency=64)
        out = {}
        for name, kid in (("spmv_a_dual", P.PM_KERNEL_SPMV_A), ("spmv_at_primal", P.PM_KERNEL_SPMV_AT)):
            ms, b = ctx.time_kernel(kid, reps)
            out[name] = {"us": ms * 1e3, "alg_bytes": b, "GBps": b / (ms * 1e-3) / 1e9,
                         "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out["variants"] = ctx.spmv_variants()
        out["lp"] = {"m": lp.m, "n": lp.n, "nnz": lp.nnz}
    return out


def bench_rolling(args, P) -> dict:
    from pramaan_amd import rolling as R
    if not WEIGHTS.exists():
        raise SystemExit(f"{WEIGHTS} is missing: train it with `python -m pramaan_amd.rolling --out {WEIGHTS}`")
    rh = P.RollingHorizon()
    nw = min(args.windows, rh.windows)
    t0 = time.perf_counter()
    lps = [rh.window(w) for w in range(nw)]
    net = R.load(WEIGHTS)
    b = P.Batch(lps, tolerance=args.tol, max_iterations=args.max_iters, check_frequency=64)
    if args.no_graph:
        b.set_graph(False)
    log(f"{nw} windows resident in {time.perf_counter() - t0:.1f}s; spmv lanes (A, A^T) {b.spmv_variants()}")
    cpp, rpp = rh.cols_per_period, rh.rows_per_period

    def clear():
        for i in range(nw):
            b.set_warm_start(i, None, None)

    def cold_step():
        clear()
        t = time.perf_counter()
        res, dev, _ = b.run()
        return time.perf_counter() - t, dev, res

    def learned_step():
        clear()
        t = time.perf_counter()
        inf = b.learned_warm_start(net, cpp, rpp)
        res, dev, _ = b.run()
        return time.perf_counter() - t, dev + inf, res, inf

    for w in range(args.warmup):
        cold_step()
        learned_step()
    cold_t, learned_t, infer_t, learned_dev = [], [], [], []
    cold_res = learned_res = None
    for s in range(args.steps):
        wall, dev, cold_res = cold_step()
        cold_t.append(wall)
        wall2, dev2, learned_res, inf = learned_step()
        learned_t.append(wall2)
        learned_dev.append(dev2)
        infer_t.append(inf)
        log(f"step {s}: cold {wall:.3f}s (device {dev:.3f}s)  learned {wall2:.3f}s (device {dev2:.3f}s, "
            f"inference {inf * 1e3:.1f} ms)")
    acc, start_kkt = b.warm_start_verdicts()
    sec_l, sec_c = float(np.mean(learned_t)), float(np.mean(cold_t))

    it_c = np.array([r.iterations for r in cold_res])
    it_l = np.array([r.iterations for r in learned_res])
    fx = json.loads((ROOT / "tests" / "golden" / "fixtures.json").read_text())["rolling"]
    obj_err = max(abs(learned_res[int(w)].objective - v) / (1 + abs(v)) for w, v in fx.items() if int(w) < nw)

    # the reference's pv / dv of every learned solution (one more learned
    # solve, untimed, that also returns the vectors)
    b.learned_warm_start(net, cpp, rpp)
    vec_res, _, _ = b.run(want_vectors=True)
    pvdv = np.array([reference_residuals(lp, r.x, r.y, args.tol) for lp, r in zip(lps, vec_res)])

    # What the network's primal-weight prediction does: the same learned
    # solve with it applied (untimed, beside the timed steps, which start
    # every window from the default weight).
    net.set_omega_head(True)
    _, dev_on, res_on, _ = learned_step()
    net.set_omega_head(False)
    it_on = np.array([r.iterations for r in res_on])
    log(f"learned start with the predicted primal weight: device {dev_on:.3f}s, iterations mean {it_on.mean():.0f}")

    kern = {}
    for name, kid in (("batch_spmv_a_dual", P.PM_KERNEL_SPMV_A), ("batch_spmv_at_primal", P.PM_KERNEL_SPMV_AT)):
        ms, by = b.time_kernel(kid, args.kernel_reps)
        kern[name] = (ms, by)
        log(f"kernel {name}: {ms * 1e3:.1f} us, {by / 1e6:.1f} MB -> {by / ms / 1e6:.0f} GB/s")
    dom = max(kern, key=lambda k: kern[k][0])
    ms, by = kern[dom]
    achieved = by / (ms * 1e-3) / 1e9
    traffic = measured_traffic(dom)
    b.close()

    out = {
        "metric": METRIC, "value": sec_l, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec_l * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (rolling-horizon blending plan, 523 periods; network trained on perturbed "
                                "plans solved by this engine)",
        "config": {"workload": f"{nw}-window rolling-horizon sequence, 24-period windows of "
                               f"{lps[0].m} rows x {lps[0].n} cols ({lps[0].nnz} nnz), solved to rel KKT "
                               f"{args.tol:g} in one batch: learned warm start (value) vs cold start",
                   "windows": nw, "tolerance": args.tol, "rows": lps[0].m, "cols": lps[0].n, "nnz": lps[0].nnz},
        "cold_seconds": sec_c, "learned_seconds": sec_l, "speedup_learned_vs_cold": sec_c / sec_l,
        "inference_ms": 1e3 * float(np.mean(infer_t)),
        "optimal": {"cold": int(sum(r.status == "OPTIMAL" for r in cold_res)),
                    "learned": int(sum(r.status == "OPTIMAL" for r in learned_res))},
        "iterations": {"cold_mean": float(it_c.mean()), "cold_max": int(it_c.max()),
                       "learned_mean": float(it_l.mean()), "learned_max": int(it_l.max())},
        "warm_starts_kept": int((acc == 1).sum()), "warm_starts_replaced_by_cold": int((acc == 0).sum()),
        "omega_head": {"learned_device_seconds_without": float(np.mean(learned_dev)),
                       "learned_device_seconds_with": dev_on,
                       "iterations_mean_without": float(it_l.mean()), "iterations_mean_with": float(it_on.mean()),
                       "note": "the learned start from the default primal weight (the timed steps) and from the "
                               "network's predicted weight"},
        "objective_rel_error_vs_highs": obj_err,
        "reference_residuals_learned": {"pv_max": float(pvdv[:, 0].max()), "dv_max": float(pvdv[:, 1].max()),
                                        "definition": "pdhg_solver.cu:107-161 (max over windows, tol = the "
                                                      "tolerance)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                     "kernel_us": ms * 1e3, "alg_bytes_per_launch": by,
                     "achieved_definition": "CSR bytes of the product and its fused update per launch / time",
                     "traffic_GBps": None if traffic is None else traffic / (ms * 1e-3) / 1e9,
                     "traffic_frac": None if traffic is None else traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic_source": "profiles/batch_iterate_rolling500_traffic.md (rocprofv3 FETCH_SIZE + "
                                       "WRITE_SIZE, calibrated on a 256 MiB copy)"},
        "batch_kernels": {k: {"us": v[0] * 1e3, "alg_bytes": v[1], "GBps": v[1] / v[0] / 1e6}
                          for k, v in kern.items()},
    }
    if args.previous > 0:
        # Previous-window warm start: inherently sequential (window w starts
        # from window w-1's solution shifted by one period), so it runs on
        # the single-LP engine; compared by iterations on the same windows.
        npw = min(args.previous, nw)
        st_p, _ = R.solve_previous(lps[:npw], cpp, rpp, args.tol)
        out["previous_window"] = {
            "windows": npw, "seconds_sequential": st_p.seconds,
            "optimal": int(sum(s == "OPTIMAL" for s in st_p.statuses)),
            "iterations_mean": float(st_p.iterations.mean()),
            "cold_iterations_mean_same_windows": float(it_c[:npw].mean()),
            "learned_iterations_mean_same_windows": float(it_l[:npw].mean())}
        log(f"previous-window warm start on windows 0..{npw - 1}: iterations mean {st_p.iterations.mean():.0f} "
            f"(cold {it_c[:npw].mean():.0f}, learned {it_l[:npw].mean():.0f}), {st_p.seconds:.2f}s sequential")
    if args.plain_windows > 0:
        out["plain_pdhg_gpu"] = plain_pdhg_gpu(args, P, lps, sec_l)
    if not args.no_lp1:
        out["spmv_lp1"] = spmv_lp1(P, args.kernel_reps)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_rolling(args, lps)
    return out


def plain_pdhg_gpu(args, P, lps, learned_seconds) -> dict:
    """Plain cold-start PDHG on the GPU: the reference algorithm
    (PM_ALG_REFERENCE: pdhg_solver.cu's iteration, restarts and pv/dv test,
    100,000-iteration budget) on the single-LP engine, run to its own end on
    a sample of windows spread over the sequence, one after another.  The
    FP64 relative KKT error of each final point is the batch engine's judge
    of that point as a warm start (pm_batch_warm_start_verdicts)."""
    k = min(len(lps), args.plain_windows)
    pick = [int(round(i * (len(lps) - 1) / max(1, k - 1))) for i in range(k)]
    res, secs = [], []
    for w in pick:
        t0 = time.perf_counter()
        r = P.solve(lps[w], algorithm="reference", tolerance=args.tol, max_iterations=100000)
        secs.append(time.perf_counter() - t0)
        res.append(r)
        log(f"plain PDHG, window {w}: {r.status} after {r.iterations} iterations, {secs[-1]:.2f}s, "
            f"pv {r.primal_residual:.2e} dv {r.dual_residual:.2e}")
    sample = [lps[w] for w in pick]
    with P.Batch(sample, tolerance=args.tol, max_iterations=64, check_frequency=64) as jb:
        for i, r in enumerate(res):
            jb.set_warm_start(i, r.x, r.y)
        jb.run()
        acc, kkt = jb.warm_start_verdicts()
    rel = [float(kk) if a == 1 else None for a, kk in zip(acc, kkt)]
    reached = sum(1 for v in rel if v is not None and v <= args.tol)
    per_window = float(np.mean(secs))
    est = per_window * len(lps)
    log(f"plain PDHG: {reached}/{k} sample windows at rel KKT <= {args.tol:g}; {per_window:.2f}s per window "
        f"-> {est:.0f}s for {len(lps)} windows in sequence")
    return {
        "sample_windows": pick, "seconds_per_window": per_window,
        "status": [r.status for r in res], "iterations": [int(r.iterations) for r in res],
        "reference_pv": [float(r.primal_residual) for r in res], "reference_dv": [float(r.dual_residual) for r in res],
        "rel_kkt_final": rel, "reached_rel_kkt_tolerance": reached,
        "seconds_all_windows_estimate": est,
        "learned_speedup_lower_bound": est / learned_seconds if reached < k else None,
        "note": ("the reference algorithm stops on its own absolute pv/dv test or at 100,000 iterations; "
                 "rel_kkt_final is the FP64 relative KKT error of its last point (None: worse than the zero "
                 "start); where no sample window reaches the tolerance, the estimate (per-window seconds x "
                 "windows) is a lower bound on its time to tolerance"),
    }


def cpu_baseline_rolling(args, lps) -> dict:
    """The reference algorithm on the CPU, measured to its own termination
    on a sample of windows (see the module docstring)."""
    nt = cpu_threads()
    k = min(len(lps), max(1, args.cpu_windows or nt))
    pick = [int(round(i * (len(lps) - 1) / max(1, k - 1))) for i in range(k)]
    wall, res = oracle_to_termination([lps[w] for w in pick], nt, args.tol)
    reached = [r for r in res if r["status"] == "OPTIMAL"]
    its = np.array([r["iterations"] for r in res])
    per_window = float(np.mean([r["thread_seconds"] for r in res]))
    waves = -(-k // nt)
    log(f"cpu baseline: {len(reached)}/{k} sample windows reached pv, dv <= {args.tol:g} "
        f"({wall:.1f}s on {nt} threads, {waves} wave(s)); {per_window:.1f} thread-s per window")
    return {
        "value": wall * len(lps) / k, "unit": "s", "cores": nt, "kind": "port",
        "sample": (f"the oracle's restated reference PDHG (FP32, pv and dv <= {args.tol:g} checked every 100 "
                   f"iterations, at most 100,000 iterations) run to its end on windows {pick}, one per host thread "
                   f"on {nt} threads at once ({wall:.2f} s wall); value = that wall time x {len(lps)}/{k}"),
        "sample_windows": k, "reached_tolerance": len(reached),
        "iterations_mean": float(its.mean()), "iterations_max": int(its.max()),
        "pv_max": float(max(r["primal_residual"] for r in res)), "dv_max": float(max(r["dual_residual"] for r in res)),
        "single_thread": {"value": per_window * len(lps), "cores": 1,
                          "sample": f"mean CPU seconds of one sample window on its own thread ({per_window:.2f} s) "
                                    f"x {len(lps)}"},
        "note": "the reference PDHG's own termination is absolute: max row/bound violation and the scaled "
                "reduced-cost violation both <= tol (pdhg_solver.cu:124-157,356)"}


def bench_batch(args, P) -> dict:
    """4,096 independent generated LPs (m log-uniform in [50, 5000], n = 2.5 m,
    4-16 nonzeros per column) in one batch, cold start to the tolerance."""
    t0 = time.perf_counter()
    lps = [P.generate(3, seed=s) for s in range(1, args.instances + 1)]
    nnz = sum(lp.nnz for lp in lps)
    log(f"generated {len(lps)} LPs, {sum(lp.m for lp in lps)} rows, {nnz} nonzeros in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    b = P.Batch(lps, tolerance=args.tol, max_iterations=args.max_iters, check_frequency=64)
    classes = b.size_classes()
    log(f"resident in {time.perf_counter() - t0:.1f}s; iterate kernels (A, A^T) {b.spmv_variants()}; "
        f"size classes {classes}")
    for _ in range(args.warmup):
        b.run()
    times, res, inst = [], None, None
    for s in range(args.steps):
        t = time.perf_counter()
        res, dev, inst = b.run()
        times.append(time.perf_counter() - t)
        log(f"step {s}: {times[-1]:.3f}s (device {dev:.3f}s)")
    sec = float(np.mean(times))
    it = np.array([r.iterations for r in res])
    err = max(abs(r.objective - lp.p_star) / (1 + abs(lp.p_star)) for r, lp in zip(res, lps))
    kern = {}
    for name, kid in (("batch_spmv_a_dual", P.PM_KERNEL_SPMV_A), ("batch_spmv_at_primal", P.PM_KERNEL_SPMV_AT)):
        kern[name] = b.time_kernel(kid, args.kernel_reps)
    dom = max(kern, key=lambda k: kern[k][0])
    ms, by = kern[dom]
    achieved = by / (ms * 1e-3) / 1e9
    b.close()
    return {
        "metric": "instances/s to 1e-6 rel KKT (batched PDHG, 1 GPU)", "value": len(lps) / sec, "unit": "instances/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(lps)} generated LPs (m log-uniform [50, 5000], n = 2.5 m, 4-16 nnz/col), "
                               f"{nnz} nonzeros, one batch, cold start to rel KKT {args.tol:g}",
                   "instances": len(lps), "nnz": nnz, "tolerance": args.tol},
        "optimal": int(sum(r.status == "OPTIMAL" for r in res)),
        "instance_seconds_p50": float(np.median(inst)), "instance_seconds_p99": float(np.quantile(inst, 0.99)),
        "iterations_mean": float(it.mean()), "iterations_max": int(it.max()),
        "objective_rel_error_vs_closed_form_max": err,
        "size_classes": classes,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom, "kernel_us": ms * 1e3,
                     "alg_bytes_per_launch": by},
    }


def bench_lp1(args, P) -> dict:
    lp = P.generate(1 if args.lp == "full" else 2, seed=1)
    if args.dense_rows > 0:
        lp = P.with_dense_rows(lp, args.dense_rows)
    ctx = P.Context(0)
    t0 = time.perf_counter()
    ctx.load(lp, algorithm="pdlp", termination="relative_kkt", tolerance=args.tol, max_iterations=args.max_iters,
             check_frequency=64, spmv_variant=args.spmv)
    load_s = time.perf_counter() - t0
    log(f"loaded m={lp.m} n={lp.n} nnz={lp.nnz} in {load_s:.2f}s (transpose on the host; scaling, "
        f"autotuning and power iteration on the device)")
    for _ in range(args.warmup):
        ctx.solve()
    times, res = [], None
    for s in range(args.steps):
        t0 = time.perf_counter()
        res = ctx.solve()
        times.append(time.perf_counter() - t0)
        log(f"step {s}: {res.status} it={res.iterations} kkt={res.rel_kkt:.2e} t={times[-1]:.3f}s")
    sec = float(np.mean(times))
    kern = {}
    for name, kid in (("spmv_a_dual", P.PM_KERNEL_SPMV_A), ("spmv_at_primal", P.PM_KERNEL_SPMV_AT)):
        kern[name] = ctx.time_kernel(kid, args.kernel_reps)
    dom = max(kern, key=lambda k: kern[k][0])
    ms, by = kern[dom]
    achieved = by / (ms * 1e-3) / 1e9
    ctx_variants = ctx.spmv_variants()
    ctx.close()
    out = {
        "metric": METRIC, "value": sec, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"single LP m={lp.m} n={lp.n} nnz={lp.nnz}, cold start to rel KKT {args.tol:g}",
                   "m": lp.m, "n": lp.n, "nnz": lp.nnz, "tolerance": args.tol},
        "status": res.status, "iterations": res.iterations, "rel_kkt": res.rel_kkt, "load_seconds": load_s,
        "reference_residuals": {"pv": res.primal_residual, "dv": res.dual_residual,
                                "definition": "pdhg_solver.cu:107-161"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom, "kernel_us": ms * 1e3,
                     "alg_bytes_per_launch": by, "spmv_variants": ctx_variants},
    }
    if not args.no_cpu:
        per_it, iters = oracle_seconds_per_iteration(lp, args.cpu_budget)
        out["cpu_baseline"] = {"value": per_it * res.iterations, "unit": "s", "cores": 1, "kind": "port",
                               "sample": f"{iters} oracle iterations on the same LP x the GPU's {res.iterations}"}
    return out


def bench_ruiz(args, P) -> dict:
    """pm_ruiz_scale (the reference's ruiz_scale, 10 passes) on the 500k x 1M
    LP.  The reference's interface scales a host-side model in place, so a
    call moves the LP to the device and back: `value` is that whole call;
    the kernels' own durations come from a rocprofv3 kernel trace of this
    command."""
    from oracle import binding as O
    lp = P.generate(1, seed=1)
    for _ in range(args.warmup):
        P.ruiz_scale(lp, 10)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        got = P.ruiz_scale(lp, 10)
        times.append(time.perf_counter() - t0)
    sec = float(np.mean(times))
    # in: row pointers, indexes and values, c, four bound arrays; out:
    # values, c, the bounds and both factor arrays
    host_bytes = (4 * (lp.m + 1) + 12 * lp.nnz + 8 * lp.n + 16 * (lp.m + lp.n)) \
        + (8 * lp.nnz + 8 * lp.n + 16 * (lp.m + lp.n) + 8 * (lp.m + lp.n))
    out = {
        "metric": "s per ruiz_scale call (10 passes, host model in and out)", "value": sec, "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Ruiz equilibration, 10 passes, of the LP m={lp.m} n={lp.n} nnz={lp.nnz}; "
                               "host arrays to the device and back inside the timed call",
                   "m": lp.m, "n": lp.n, "nnz": lp.nnz, "passes": 10},
        "host_device_bytes": host_bytes,
        "roofline": None,
        "roofline_note": "kernel durations per pass: profiles/ruiz_scale_lp1_kernel_trace.md",
    }
    if not args.no_cpu:
        t0 = time.perf_counter()
        want = O.ruiz_scale(lp, 10)
        cpu = time.perf_counter() - t0
        out["bitwise_equal_to_oracle"] = all(np.array_equal(got[k], want[k]) for k in want)
        out["cpu_baseline"] = {"value": cpu, "unit": "s", "cores": 1, "kind": "port",
                               "sample": "one full call of the oracle's restated ruiz_scale on the same LP"}
    return out


def bench_traffic(args, P) -> dict:
    """Dispatches for the rocprofv3 counter runs of the batch iterate kernels:
    the two iterate kernels on the rolling-horizon batch with every window
    running, then a plain device copy of 256 MiB that calibrates the
    counters (bytes = counter x 1024 x copy bytes / the counter on the copy).
    The batch's autotuner also runs every candidate kernel at creation."""
    import torch
    rh = P.RollingHorizon()
    lps = [rh.window(w) for w in range(min(args.windows, rh.windows))]
    out = {"workload": "traffic", "windows": len(lps), "kernels": {}}
    with P.Batch(lps, tolerance=args.tol, check_frequency=64) as b:
        out["variants"] = b.spmv_variants()
        log(f"iterate kernels (A, A^T): {out['variants']}")
        for which, name in ((P.PM_KERNEL_SPMV_AT, "batch_spmv_at_primal"), (P.PM_KERNEL_SPMV_A, "batch_spmv_a_dual")):
            ms, by = b.time_kernel(which, args.kernel_reps)
            out["kernels"][name] = {"us": ms * 1e3, "alg_bytes": by}
            log(f"{name}: {ms * 1e3:.1f} us, algorithmic {by / 1e6:.1f} MB")
    src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_bytes"] = src.numel() * 4
    log(f"copy: {out['copy_bytes']} bytes")
    return out


def bench_train(args, P) -> dict:
    """Training of the warm-start network (hand-written HIP forward, backward
    and Adam) on 20,000 rolling-horizon windows: 40 perturbed plans of 500
    windows (demands x U(0.8, 1.2), blend coefficients x U(0.9, 1.1)) solved
    cold by the batch engine.  A step is one epoch over the training set
    (mini-batches of 16 windows, graphs rebuilt from each window); `value` is
    samples per second of the timed epochs by the wall clock (host threads
    building the next mini-batch's graph while the device trains on the
    current one); `samples_per_s_device` counts only the graph uploads and
    the device steps."""
    from pramaan_amd import rolling as R
    t0 = time.perf_counter()
    ts = R.training_set(range(1, args.plans + 1), 500, log=lambda m: None)
    build_s = time.perf_counter() - t0
    log(f"training set: {len(ts)} windows from {args.plans} plans in {build_s:.1f}s")
    net = P.Gnn(seed=1)
    if args.warmup:
        R.train(net, ts, epochs=args.warmup, log=lambda m: None)
    t0 = time.perf_counter()
    tr = R.train(net, ts, epochs=args.steps, log=lambda m: log(m))
    wall = time.perf_counter() - t0
    hist = tr["history"]
    return {
        "metric": "training samples/s of the warm-start network (HIP forward + backward + Adam, 1 GPU)",
        "value": tr["samples_per_s"], "samples_per_s_device": tr["samples_per_s_device"],
        "unit": "samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / max(1, args.steps) * 1e3, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "bf16 MFMA (FP32 accumulate)", "data": "synthetic (solver-produced targets)",
        "config": {"workload": f"{len(ts)} rolling-horizon windows ({args.plans} perturbed plans x 500), "
                               "one epoch per step, mini-batches of 16 windows", "samples": len(ts),
                   "training_set_seconds": build_s},
        "loss": {"train": hist[-1]["train_loss"], "validation": hist[-1]["val_loss"],
                 "zero_predictor": hist[-1]["zero_predictor_loss"]},
        "roofline": None,
    }


def bench_starts(args, P) -> dict:
    """What a warm start can give on the rolling windows: one resident batch of
    the windows, solved to the tolerance from starts built out of each
    window's own cold solution (perturbed, or mixed with the network's
    prediction, or snapped onto its true active set: columns at their
    bound, zero duals on slack rows).  One run per start; the first (cold)
    also pays the graph capture.  Each line: device seconds, mean and max
    iterations, windows optimal."""
    from pramaan_amd import rolling as R
    from pramaan_amd import warmstart as W
    rh = P.RollingHorizon()
    nw = min(args.windows, rh.windows)
    lps = [rh.window(w) for w in range(nw)]
    graphs = [W.graph_of(lp, rh.cols_per_period, rh.rows_per_period) for lp in lps]
    net = R.load(WEIGHTS)
    net_starts, _ = R.predict(net, lps, graphs)
    rng = np.random.default_rng(0)
    out = {}
    with P.Batch(lps, tolerance=args.tol, max_iterations=args.max_iters, check_frequency=64) as b:
        def run(name, st=None, want=False):
            for i in range(nw):
                b.set_warm_start(i, *(st[i] if st is not None else (None, None)))
            res, secs, _ = b.run(want_vectors=want)
            it = np.array([r.iterations for r in res])
            out[name] = {"seconds": secs, "iterations_mean": float(it.mean()), "iterations_max": int(it.max()),
                         "optimal": int(sum(r.status == "OPTIMAL" for r in res))}
            log(f"{name}: {json.dumps(out[name])}")
            return res
        cold = run("cold", want=True)
        xs, ys = [r.x.copy() for r in cold], [r.y.copy() for r in cold]
        run("learned", net_starts)
        run("exact", list(zip(xs, ys)))
        for nz in (0.01, 0.1, 0.3):
            run(f"exact_noise_{nz}", [(np.clip(x * (1 + nz * rng.uniform(-1, 1, len(x))), lp.var_lower,
                                               lp.var_upper), y * (1 + nz * rng.uniform(-1, 1, len(y))))
                                      for x, y, lp in zip(xs, ys, lps)])
        run("true_x_network_y", [(x, s[1]) for s, x in zip(net_starts, xs)])
        run("network_x_true_y", [(s[0], y) for s, y in zip(net_starts, ys)])
        classes, snapped = [], []
        for (x0, y0), x, y, lp in zip(net_starts, xs, ys, lps):
            atl = np.abs(x - lp.var_lower) <= 1e-7 * (1 + np.abs(lp.var_lower))
            atu = np.abs(x - lp.var_upper) <= 1e-7 * (1 + np.abs(lp.var_upper))
            cl = np.where(atl, 0, np.where(atu, 2, 1))
            classes.append(cl)
            snapped.append((np.where(atl, lp.var_lower, np.where(atu, lp.var_upper, x0)),
                            np.where(np.abs(y) <= 1e-9 * (1 + np.abs(y).max()), 0.0, y0)))
        run("network_on_true_active_set", snapped)
        run("network_x_on_true_active_set", [(s[0], n[1]) for s, n in zip(snapped, net_starts)])
        for q in (0.05, 0.1):
            st = []
            for (x0, y0), cl, lp in zip(net_starts, classes, lps):
                c2 = np.where(rng.random(len(cl)) < q, (cl + rng.integers(1, 3, len(cl))) % 3, cl)
                lo = np.where(np.abs(lp.var_lower) < 1e29, lp.var_lower, x0)
                hi = np.where(np.abs(lp.var_upper) < 1e29, lp.var_upper, x0)
                st.append((np.where(c2 == 0, lo, np.where(c2 == 2, hi, x0)), y0))
            run(f"network_x_on_active_set_accuracy_{1 - q:.2f}", st)
    allc = np.concatenate(classes)
    return {
        "metric": "device s to rel KKT, rolling-horizon batch, by warm start", "value": out["learned"]["seconds"],
        "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": out["learned"]["seconds"] * 1e3,
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nw} rolling-horizon windows, one batch, one solve per start", "windows": nw,
                   "tolerance": args.tol},
        "column_classes": {"at_lower": float((allc == 0).mean()), "interior": float((allc == 1).mean()),
                           "at_upper": float((allc == 2).mean())},
        "starts": out, "roofline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["rolling500", "lp1", "batch4096", "ruiz", "train", "traffic", "starts"],
                    default="rolling500")
    ap.add_argument("--plans", type=int, default=40, help="training plans of the train workload (500 windows each)")
    ap.add_argument("--cpu-windows", type=int, default=0,
                    help="windows of the CPU baseline sample (default: one per host thread)")
    ap.add_argument("--instances", type=int, default=4096, help="LPs of the batch4096 workload")
    ap.add_argument("--windows", type=int, default=500)
    ap.add_argument("--previous", type=int, default=100,
                    help="windows of the sequential previous-window warm-start comparison (0 skips it)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iters", type=int, default=200000)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--spmv", default="auto", help="A x-bar / A^T y kernel of the lp1 workload (a name of _capi.SPMV)")
    ap.add_argument("--lp", choices=["full", "tenth"], default="full",
                    help="lp1 workload: the 500k x 1M LP or its 50k x 100k tenth")
    ap.add_argument("--dense-rows", type=int, default=0,
                    help="lp1 workload: rows over every column appended (skewed row lengths)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--plain-windows", type=int, default=4,
                    help="sample windows of the plain cold-start GPU PDHG comparison (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lp1", action="store_true", help="skip the 500k x 1M SpMV measurement")
    ap.add_argument("--no-graph", action="store_true",
                    help="issue the batch launches directly instead of replaying the HIP graph (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs every LP of the workload on a single device and has no "
              "multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    from pramaan_amd import _capi as P
    if args.workload == "rolling500":
        out = bench_rolling(args, P)
    elif args.workload == "batch4096":
        out = bench_batch(args, P)
    elif args.workload == "ruiz":
        out = bench_ruiz(args, P)
    elif args.workload == "train":
        out = bench_train(args, P)
    elif args.workload == "traffic":
        out = bench_traffic(args, P)
    elif args.workload == "starts":
        out = bench_starts(args, P)
    else:
        out = bench_lp1(args, P)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
