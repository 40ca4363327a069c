"""Rolling-horizon solving on the GPU: cold start, previous-window warm start
and learned warm start (pm_gnn_* network) for the windows of a
multi-period blending plan (pm_rh_*).

* cold:      every window in one batched, graph-captured launch from zero.
* previous:  windows one after another on the single-LP engine, window w
             starting from window w-1's solution shifted by one period
             (columns/rows of periods 1..23 move to 0..22, the last period
             repeats) -- inherently sequential.
* learned:   the network predicts x, y for every window at once, then all
             windows are solved in one batched launch from those points.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import _capi as P
from . import warmstart as W


def shift(v: np.ndarray, per: int) -> np.ndarray:
    return np.concatenate([v[per:], v[-per:]])


@dataclass
class RunStats:
    seconds: float            # device-side time of the solves (and inference)
    iterations: np.ndarray
    statuses: list
    objectives: np.ndarray
    rel_kkt: np.ndarray
    instance_seconds: np.ndarray | None = None
    inference_seconds: float = 0.0


def _stats(res, seconds, inst=None, infer=0.0):
    return RunStats(seconds, np.array([r.iterations for r in res]), [r.status for r in res],
                    np.array([r.objective for r in res]), np.array([r.rel_kkt for r in res]), inst, infer)


def solve_cold(lps, tol, max_iterations=200000, want_vectors=False):
    with P.Batch(lps, tolerance=tol, max_iterations=max_iterations, check_frequency=64) as b:
        res, secs, inst = b.run(want_vectors=want_vectors)
    return _stats(res, secs, inst), res


def solve_previous(lps, tol, cols_pp, rows_pp, max_iterations=200000):
    ctx = P.Context(0)
    res, secs = [], 0.0
    x0 = y0 = None
    for lp in lps:
        kw = dict(algorithm="pdlp", termination="relative_kkt", tolerance=tol, max_iterations=max_iterations,
                  check_frequency=64)
        if x0 is not None:
            kw.update(x0=shift(x0, cols_pp), y0=shift(y0, rows_pp))
        ctx.load(lp, **kw)
        t0 = time.perf_counter()
        r = ctx.solve(want_vectors=True)
        secs += time.perf_counter() - t0
        res.append(r)
        x0, y0 = r.x, r.y
    ctx.close()
    return _stats(res, secs), res


def predict(net: "P.Gnn", lps, graphs, chunk=32):
    """Warm starts for every window; returns (list of (x0, y0), device seconds)."""
    out, secs = [], 0.0
    for s in range(0, len(lps), chunk):
        gs = graphs[s:s + chunk]
        bg = W.batch_graphs(gs)
        net.set_graph(bg)
        t0 = time.perf_counter()
        xh, yh = net.forward()
        secs += time.perf_counter() - t0
        xo = yo = 0
        for lp, g in zip(lps[s:s + chunk], gs):
            out.append(W.warm_start(lp, g, xh[xo:xo + g["n"]], yh[yo:yo + g["m"]]))
            xo += g["n"]
            yo += g["m"]
    return out, secs


def solve_learned(lps, graphs, net, tol, max_iterations=200000):
    starts, infer = predict(net, lps, graphs)
    with P.Batch(lps, tolerance=tol, max_iterations=max_iterations, check_frequency=64) as b:
        for i, (x0, y0) in enumerate(starts):
            b.set_warm_start(i, x0, y0)
        res, secs, inst = b.run()
    return _stats(res, secs + infer, inst, infer), res


def training_data(seeds, windows_per_seed, tol=1e-5, demand_noise=0.2, blend_noise=0.1, log=print):
    """(graph, tx, ty) tuples from perturbed plans solved by the batch engine."""
    data = []
    for sd in seeds:
        rh = P.RollingHorizon(seed=1000 + sd, demand_noise=demand_noise, blend_noise=blend_noise,
                              perturb_seed=sd)
        lps = [rh.window(w) for w in range(min(windows_per_seed, rh.windows))]
        st, res = solve_cold(lps, tol, want_vectors=True)
        ok = 0
        for lp, r in zip(lps, res):
            if r.status != "OPTIMAL":
                continue
            g = W.graph_of(lp, rh.cols_per_period, rh.rows_per_period)
            tx, ty = W.targets(lp, g, r.x, r.y)
            data.append((g, tx.astype(np.float32), ty.astype(np.float32)))
            ok += 1
        log(f"seed {sd}: {ok}/{len(lps)} windows solved in {st.seconds:.2f}s (iters mean {st.iterations.mean():.0f})")
    return data


def train(net, data, epochs=2, batch=8, lr=2e-3, log=print, seed=0):
    """Adam over mini-batches of windows; returns (samples/s, final loss)."""
    rng = np.random.default_rng(seed)
    n = len(data)
    t0 = time.perf_counter()
    seen = 0
    loss = float("nan")
    for ep in range(epochs):
        order = rng.permutation(n)
        losses = []
        for s in range(0, n, batch):
            idx = order[s:s + batch]
            bg = W.batch_graphs([data[i][0] for i in idx])
            tx = np.concatenate([data[i][1] for i in idx])
            ty = np.concatenate([data[i][2] for i in idx])
            net.set_graph(bg)
            step_lr = lr * (0.5 * (1 + np.cos(np.pi * (ep * n + s) / (epochs * n))))
            losses.append(net.train_step(tx, ty, lr=float(step_lr)))
            seen += len(idx)
        loss = float(np.mean(losses[-max(1, len(losses) // 10):]))
        log(f"epoch {ep}: loss {loss:.4e} ({seen / (time.perf_counter() - t0):.1f} samples/s)")
    return seen / (time.perf_counter() - t0), loss
