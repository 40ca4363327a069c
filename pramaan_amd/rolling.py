"""Rolling-horizon solving on the GPU: cold start, previous-window warm start
and learned warm start (pm_gnn_* network) for the windows of a
multi-period blending plan (pm_rh_*), and the training of that network on
solutions the batch engine produces.

* cold:      every window in one batched, graph-captured launch from zero.
* previous:  windows one after another on the single-LP engine, window w
             starting from window w-1's solution shifted by one period
             (columns / rows of periods 1..23 move to 0..22, the last
             period repeats) -- inherently sequential.
* learned:   the network predicts x, y for every window at once, then all
             windows are solved in one batched launch from those points.
             Every prediction is judged by its FP64 relative KKT error
             (the engine's own check): a window whose prediction is worse
             than the cold start's zero point falls back to cold.

Training data are (window, x, y) from perturbed plans (demand and blend
coefficients redrawn under new seeds) solved cold by the batch engine; the
windows are regenerated from their plan when a mini-batch is built, so only
the targets are stored.

    python -m pramaan_amd.rolling --plans 40 --epochs 8 --out weights.npz
"""
from __future__ import annotations

import argparse
import json
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _capi as P
from . import warmstart as W

TOL = 1e-6
HELD_OUT = 1000  # perturbation seeds from here on never enter a training set


def shift(v: np.ndarray, per: int) -> np.ndarray:
    """A window's vector moved one period earlier, last period repeated."""
    return np.concatenate([v[per:], v[-per:]])


@dataclass
class RunStats:
    seconds: float            # device time of the solves (plus inference for "learned")
    iterations: np.ndarray
    statuses: list
    objectives: np.ndarray
    rel_kkt: np.ndarray
    instance_seconds: np.ndarray | None = None
    inference_seconds: float = 0.0
    fallbacks: int = 0

    def summary(self) -> dict:
        return dict(seconds=self.seconds, iterations_mean=float(self.iterations.mean()),
                    iterations_p50=float(np.median(self.iterations)), iterations_max=int(self.iterations.max()),
                    optimal=int(sum(s == "OPTIMAL" for s in self.statuses)), windows=len(self.statuses),
                    inference_seconds=self.inference_seconds, fallbacks=self.fallbacks)


def _stats(res, seconds, inst=None, infer=0.0, fallbacks=0):
    return RunStats(seconds, np.array([r.iterations for r in res]), [r.status for r in res],
                    np.array([r.objective for r in res]), np.array([r.rel_kkt for r in res]), inst, infer, fallbacks)


def solve_cold(lps, tol=TOL, max_iterations=200000, want_vectors=False, batch=None):
    own = batch is None
    b = batch or P.Batch(lps, tolerance=tol, max_iterations=max_iterations, check_frequency=64)
    try:
        for i in range(len(lps)):
            b.set_warm_start(i, None, None)
        res, secs, inst = b.run(want_vectors=want_vectors)
    finally:
        if own:
            b.close()
    return _stats(res, secs, inst), res


def solve_previous(lps, cols_pp, rows_pp, tol=TOL, max_iterations=200000):
    """Sequential previous-window warm start on the single-LP engine."""
    ctx = P.Context(0)
    res, secs = [], 0.0
    x0 = y0 = None
    try:
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
    finally:
        ctx.close()
    return _stats(res, secs), res


def predict(net: P.Gnn, lps, graphs, chunk=32):
    """Warm starts for every window: list of (x0, y0), and the seconds spent
    in the network (graph upload and forward)."""
    out, secs = [], 0.0
    for s in range(0, len(lps), chunk):
        gs = graphs[s:s + chunk]
        bg = W.batch_graphs(gs)
        t0 = time.perf_counter()
        net.set_graph(bg)
        xh, yh = net.forward()
        secs += time.perf_counter() - t0
        xo = yo = 0
        for lp, g in zip(lps[s:s + chunk], gs):
            out.append(W.warm_start(lp, g, xh[xo:xo + g["n"]], yh[yo:yo + g["m"]]))
            xo += g["n"]
            yo += g["m"]
    return out, secs


def solve_learned(lps, graphs, net, tol=TOL, max_iterations=200000, batch=None):
    """Learned warm start for every window, one batched solve.  The engine
    judges each prediction by its FP64 relative KKT error against the cold
    start's and starts from the better one (pm_batch_run)."""
    starts, infer = predict(net, lps, graphs)
    own = batch is None
    b = batch or P.Batch(lps, tolerance=tol, max_iterations=max_iterations, check_frequency=64)
    try:
        for i, (x0, y0) in enumerate(starts):
            b.set_warm_start(i, x0, y0)
        res, secs, inst = b.run()
        acc, _ = b.warm_start_verdicts()
        for i in range(len(lps)):
            b.set_warm_start(i, None, None)
    finally:
        if own:
            b.close()
    return _stats(res, secs + infer, inst, infer, int((acc == 0).sum())), res


# ---------------------------------------------------------------------------
# training data
# ---------------------------------------------------------------------------

@dataclass
class TrainingSet:
    plans: list = field(default_factory=list)        # RollingHorizon objects
    index: list = field(default_factory=list)        # (plan, window)
    tx: list = field(default_factory=list)           # x / s, float32
    ty: list = field(default_factory=list)           # y / cs, float32
    omega: list = field(default_factory=list)        # final primal weight of the cold solve
    tw: list = field(default_factory=list)           # log(final / default primal weight)

    def __len__(self):
        return len(self.index)


def training_set(seeds, windows_per_plan=500, tol=1e-5, demand_noise=0.2, blend_noise=0.1, base_seed=1, log=print):
    """Solves the windows of perturbed plans cold in batches and keeps the
    optimal ones as targets.  Plan s is the base plan (seed base_seed, the
    measured sequence's) with every demand scaled by U(1 -+ demand_noise)
    and every blend coefficient by U(1 -+ blend_noise) under perturbation
    seed s; seeds >= HELD_OUT are kept for evaluation."""
    ts = TrainingSet()
    for sd in seeds:
        rh = P.RollingHorizon(seed=base_seed, demand_noise=demand_noise, blend_noise=blend_noise, perturb_seed=sd)
        pi = len(ts.plans)
        ts.plans.append(rh)
        nw = min(windows_per_plan, rh.windows)
        lps = [rh.window(w) for w in range(nw)]
        with P.Batch(lps, tolerance=tol, max_iterations=200000, check_frequency=64) as b:
            st, res = solve_cold(lps, tol, want_vectors=True, batch=b)
            scal = [b.scaling(w) for w in range(nw)]
        ok = 0
        for w, (lp, r) in enumerate(zip(lps, res)):
            if r.status != "OPTIMAL":
                continue
            s, cs = _scales(lp)
            ts.index.append((pi, w))
            ts.tx.append((r.x / s).astype(np.float32))
            ts.ty.append((r.y / cs).astype(np.float32))
            ts.omega.append(r.primal_weight)
            ts.tw.append(float(np.log(r.primal_weight / default_primal_weight(lp, *scal[w]))))
            ok += 1
        log(f"plan {sd}: {ok}/{nw} windows solved in {st.seconds:.2f}s (iterations mean {st.iterations.mean():.0f})")
    return ts


def default_primal_weight(lp, row_scale, col_scale):
    """The batch engine's initial primal weight of an LP: ||D_c c|| / ||D_r b||
    on the scaled problem (c in minimisation form, b_i the larger finite
    |bound| of row i)."""
    c = (-1.0 if lp.sense == 1 else 1.0) * lp.c
    lo = np.where(lp.row_lower > -1e30, np.abs(lp.row_lower), 0.0)
    hi = np.where(lp.row_upper < 1e30, np.abs(lp.row_upper), 0.0)
    cn = np.linalg.norm(col_scale * c)
    bn = np.linalg.norm(row_scale * np.maximum(lo, hi))
    return cn / bn if cn > 1e-10 and bn > 1e-10 else 1.0


def _scales(lp):
    b = np.concatenate([lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper])
    b = np.abs(b[np.abs(b) < 1e29])
    s = max(1.0, float(b.max()) if len(b) else 1.0)
    cs = max(1e-12, float(np.abs(lp.c).max())) if lp.n else 1.0
    return s, cs


def _minibatch(ts, idx, pool, arena=None):
    """Block-diagonal graph of the windows idx (built in place, in
    parallel, in `arena`'s page-locked memory when given), their x and y
    targets, each window's column offset and primal-weight target."""
    lps = list(pool.map(lambda k: ts.plans[ts.index[k][0]].window(ts.index[k][1]), idx))
    rh = ts.plans[ts.index[idx[0]][0]]  # every plan shares the period layout
    bg = W.batch_graphs_of(lps, rh.cols_per_period, rh.rows_per_period, pool, arena)
    bg["seg"] = np.concatenate([[0], np.cumsum([lp.n for lp in lps])]).astype(np.int32)
    bg["tw"] = np.array([ts.tw[k] for k in idx], np.float32)
    return bg, np.concatenate([ts.tx[k] for k in idx]), np.concatenate([ts.ty[k] for k in idx])


def _load(net, bg):
    net.set_graph(bg)
    net.set_segments(bg["seg"])
    net.set_omega_targets(bg["tw"])


def evaluate_loss(net, ts, idx, batch, pool):
    tot, cnt, base = 0.0, 0, 0.0
    for s in range(0, len(idx), batch):
        sel = idx[s:s + batch]
        bg, tx, ty = _minibatch(ts, sel, pool)
        _load(net, bg)
        loss, _ = net.gradient(tx, ty)
        tot += loss * len(sel)
        base += (float(np.mean(tx.astype(np.float64) ** 2)) + float(np.mean(ty.astype(np.float64) ** 2))
                 + 0.01 * float(np.mean(bg["tw"].astype(np.float64) ** 2))) * len(sel)
        cnt += len(sel)
    return tot / max(cnt, 1), base / max(cnt, 1)


def train(net, ts, epochs=8, batch=16, lr=2e-3, val_frac=0.05, log=print, seed=0, threads=16):
    """Adam over mini-batches of windows with a cosine learning rate;
    returns a dict with samples/s (wall clock of the training passes,
    mini-batch building included; samples_per_s_device counts only the
    graph uploads and the device steps) and train / validation /
    zero-predictor losses."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(ts))
    nval = max(1, int(len(ts) * val_frac))
    val, tr = order[:nval], order[nval:]
    steps_total = epochs * ((len(tr) + batch - 1) // batch)
    step = 0
    seen = 0
    t_train = 0.0
    hist = []
    # One thread assembles the next mini-batch (its windows built by the
    # pool) while the GPU trains on the current one; the two alternate
    # between two page-locked arenas (a graph's upload has finished when
    # pm_gnn_set_graph returns, before its arena is built into again).
    t_wall = 0.0
    arenas = [P.HostArena(), P.HostArena()]
    with ThreadPoolExecutor(threads) as pool, ThreadPoolExecutor(1) as ahead:
        for ep in range(epochs):
            perm = rng.permutation(tr)
            losses = []
            t_ep = time.perf_counter()
            fut = ahead.submit(_minibatch, ts, perm[:batch], pool, arenas[0]) if len(perm) else None
            for s in range(0, len(perm), batch):
                bg, tx, ty = fut.result()
                nxt = perm[s + batch:s + 2 * batch]
                fut = ahead.submit(_minibatch, ts, nxt, pool, arenas[(s // batch + 1) % 2]) if len(nxt) else None
                step_lr = lr * 0.5 * (1.0 + np.cos(np.pi * step / max(1, steps_total)))
                t0 = time.perf_counter()
                _load(net, bg)
                losses.append(net.train_step(tx, ty, lr=float(step_lr)))
                t_train += time.perf_counter() - t0
                seen += len(perm[s:s + batch])
                step += 1
            t_wall += time.perf_counter() - t_ep
            vloss, vbase = evaluate_loss(net, ts, val, batch, pool)
            rec = dict(epoch=ep, train_loss=float(np.mean(losses[-max(1, len(losses) // 5):])), val_loss=vloss,
                       zero_predictor_loss=vbase, samples_per_s=seen / max(t_wall, 1e-9),
                       samples_per_s_device=seen / max(t_train, 1e-9))
            hist.append(rec)
            log(json.dumps(rec))
    return dict(samples_per_s=seen / max(t_wall, 1e-9), samples_per_s_device=seen / max(t_train, 1e-9),
                history=hist)


def default_weights() -> Path:
    """The shipped trained parameters (trained by `python -m pramaan_amd.rolling`)."""
    return Path(__file__).resolve().parent / "data" / "warmstart_gnn.npz"


def save(net: P.Gnn, path, meta: dict | None = None) -> None:
    np.savez(path, params=net.get_params(), meta=json.dumps(meta or {}))


def load(path, device=0) -> P.Gnn:
    with np.load(path, allow_pickle=False) as z:
        net = P.Gnn(device=device)
        net.set_params(z["params"])
    return net


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--plans", type=int, default=40, help="perturbed training plans (500 windows each)")
    ap.add_argument("--windows", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=None, help="write the trained parameters (.npz)")
    ap.add_argument("--no-eval", action="store_true")
    ap.add_argument("--previous", type=int, default=0, help="windows for the sequential previous-window run")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    t0 = time.perf_counter()
    ts = training_set(range(1, a.plans + 1), a.windows, log=log)
    log(f"training set: {len(ts)} windows in {time.perf_counter() - t0:.1f}s")
    net = P.Gnn(seed=a.seed)
    tr = train(net, ts, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, seed=a.seed)
    if a.out:
        save(net, a.out, dict(plans=a.plans, windows=a.windows, epochs=a.epochs, batch=a.batch, lr=a.lr,
                              seed=a.seed, samples=len(ts)))
    if a.no_eval:
        return 0
    out = dict(training=dict(samples=len(ts), samples_per_s=tr["samples_per_s"], history=tr["history"]))
    out.update(evaluate(net, previous=a.previous, log=log))
    log(json.dumps(out))
    return 0


def evaluate(net, nwin=500, tol=TOL, previous=0, log=print) -> dict:
    """Cold vs learned (and optionally sequential previous-window) on the
    unperturbed 500-window sequence, and cold vs learned on a perturbed
    sequence whose perturbation seed no training set uses."""
    out = {}
    for name, kw in (("nominal", {}), ("held_out", dict(demand_noise=0.2, blend_noise=0.1, perturb_seed=HELD_OUT))):
        rh = P.RollingHorizon(**kw)
        lps = [rh.window(w) for w in range(min(nwin, rh.windows))]
        with P.Batch(lps, tolerance=tol, max_iterations=200000, check_frequency=64) as b:
            cold, _ = solve_cold(lps, tol, batch=b)
            for i in range(len(lps)):
                b.set_warm_start(i, None, None)
            infer = b.learned_warm_start(net, rh.cols_per_period, rh.rows_per_period)
            res, secs, inst = b.run()
            acc, _ = b.warm_start_verdicts()
            learned = _stats(res, secs + infer, inst, infer, int((acc == 0).sum()))
            net.set_omega_head(True)  # the same start from the predicted primal weight
            infer = b.learned_warm_start(net, rh.cols_per_period, rh.rows_per_period)
            res, secs, inst = b.run()
            net.set_omega_head(False)
            with_head = _stats(res, secs + infer, inst, infer)
        out[name] = dict(cold=cold.summary(), learned=learned.summary(), learned_with_omega_head=with_head.summary(),
                         speedup_learned_vs_cold=cold.seconds / learned.seconds)
        if previous and name == "nominal":
            prev, _ = solve_previous(lps[:previous], rh.cols_per_period, rh.rows_per_period, tol)
            out[name]["previous"] = prev.summary()
    log(json.dumps(out))
    return out


if __name__ == "__main__":
    raise SystemExit(main())
