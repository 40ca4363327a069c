"""Learned warm start for rolling-horizon LPs: graph features, targets, and
the mapping between network outputs and PDHG starting points.

The network itself (pm_gnn_*, HIP) sees one LP as a bipartite graph:
  column features (16): c_j / max|c|, l_j / s, u_j / s (0 when infinite),
                       1[l finite], 1[u finite], nnz_j / max nnz,
                       position of the column's period in the window,
                       role bits
  row features (16):   l_i / s, u_i / s (0 when infinite), 1[l finite],
                       1[u finite], 1[equality], nnz_i / max nnz,
                       period position, role bits
with s = max(1, largest finite |bound|).  The role bits are the 9 bits of
the node's index within its period as -1 / +1 (0 without a period): every
window of a plan shares its period layout, so they say which component,
product or row kind a node is.  Targets are x_j / s and
y_i / max|c| (y in the engine's sign).  Several LPs are batched as one
block-diagonal graph.
"""
from __future__ import annotations

import numpy as np

from ._capi import INF, LP

FEAT = 16


def graph_of(lp: LP, period_cols: int | None = None, period_rows: int | None = None) -> dict:
    """Graph arrays of one LP (see the module docstring), built by
    pm_gnn_lp_features."""
    from . import _capi as P
    m, n, nnz = lp.m, lp.n, lp.nnz
    cp = np.zeros(n + 1, np.int32)
    ri = np.zeros(max(nnz, 1), np.int32)
    ar = np.zeros(max(nnz, 1), np.float32)
    ac = np.zeros(max(nnz, 1), np.float32)
    xr = np.zeros((m, FEAT), np.float32)
    xc = np.zeros((n, FEAT), np.float32)
    sc = np.zeros(2)
    s = lp.struct()
    P._check(P.lib().pm_gnn_lp_features(P.C.byref(s), period_cols or 0, period_rows or 0, P._p(cp, P.C.c_int32),
                                        P._p(ri, P.C.c_int32), P._p(ar, P.C.c_float), P._p(ac, P.C.c_float),
                                        P._p(xr, P.C.c_float), P._p(xc, P.C.c_float), P._p(sc, P.C.c_double)))
    return dict(m=m, n=n, rp=lp.row_ptr, ci=lp.col_idx, ar=ar[:nnz], cp=cp, ri=ri[:nnz], ac=ac[:nnz], xr=xr, xc=xc,
                s=float(sc[0]), cs=float(sc[1]))


def batch_graphs(graphs: list[dict]) -> dict:
    """Block-diagonal union of several graphs."""
    out = {k: [] for k in ("rp", "ci", "ar", "cp", "ri", "ac", "xr", "xc")}
    mo = no = zo = zo2 = 0
    for g in graphs:
        out["rp"].append(g["rp"][:-1] + zo)
        out["cp"].append(g["cp"][:-1] + zo2)
        out["ci"].append(g["ci"] + no)
        out["ri"].append(g["ri"] + mo)
        for k in ("ar", "ac", "xr", "xc"):
            out[k].append(g[k])
        mo += g["m"]
        no += g["n"]
        zo += len(g["ci"])
        zo2 += len(g["ri"])
    res = {k: np.concatenate(v) for k, v in out.items()}
    res["rp"] = np.concatenate([res["rp"], [zo]])
    res["cp"] = np.concatenate([res["cp"], [zo2]])
    res["m"], res["n"] = mo, no
    return res


def batch_graphs_of(lps: list, period_cols: int | None = None, period_rows: int | None = None,
                    pool=None, arena=None) -> dict:
    """batch_graphs([graph_of(lp, ...) for lp in lps]) built in place: every
    LP's graph is written by pm_gnn_lp_features straight into its slices of
    the combined arrays, offsets added there, so nothing is concatenated
    afterwards; `pool` (a thread pool) builds the LPs' slices in parallel;
    `arena` (a _capi.HostArena) places the arrays in page-locked memory."""
    from . import _capi as P
    ms = np.array([lp.m for lp in lps], np.int64)
    ns = np.array([lp.n for lp in lps], np.int64)
    zs = np.array([lp.nnz for lp in lps], np.int64)
    mo = np.concatenate([[0], np.cumsum(ms)])
    no = np.concatenate([[0], np.cumsum(ns)])
    zo = np.concatenate([[0], np.cumsum(zs)])
    M, N, Z = int(mo[-1]), int(no[-1]), int(zo[-1])
    new = (lambda name, shape, dt: arena.array(name, shape, dt)) if arena is not None else \
        (lambda name, shape, dt: np.empty(shape, dt))
    rp = new("rp", (M + 1,), np.int32)
    cp = new("cp", (N + 1,), np.int32)
    ci = new("ci", (Z,), np.int32)
    ri = new("ri", (Z,), np.int32)
    ar = new("ar", (Z,), np.float32)
    ac = new("ac", (Z,), np.float32)
    xr = new("xr", (M, FEAT), np.float32)
    xc = new("xc", (N, FEAT), np.float32)

    def one(k):
        lp = lps[k]
        m, n, nnz = lp.m, lp.n, lp.nnz
        z0, r0, c0 = int(zo[k]), int(mo[k]), int(no[k])
        cpk = np.zeros(n + 1, np.int32)
        sc = np.zeros(2)
        pad_i = np.zeros(1, np.int32) if nnz == 0 else None
        pad_f = np.zeros(1, np.float32) if nnz == 0 else None
        s = lp.struct()
        P._check(P.lib().pm_gnn_lp_features(
            P.C.byref(s), period_cols or 0, period_rows or 0, P._p(cpk, P.C.c_int32),
            P._p(ri[z0:] if nnz else pad_i, P.C.c_int32), P._p(ar[z0:] if nnz else pad_f, P.C.c_float),
            P._p(ac[z0:] if nnz else pad_f, P.C.c_float), P._p(xr[r0:] if m else np.zeros((1, FEAT), np.float32),
                                                                P.C.c_float),
            P._p(xc[c0:], P.C.c_float), P._p(sc, P.C.c_double)))
        ri[z0:z0 + nnz] += r0
        ci[z0:z0 + nnz] = lp.col_idx
        ci[z0:z0 + nnz] += c0
        rp[r0:r0 + m] = lp.row_ptr[:-1]
        rp[r0:r0 + m] += z0
        cp[c0:c0 + n] = cpk[:-1]
        cp[c0:c0 + n] += z0

    if pool is None:
        for k in range(len(lps)):
            one(k)
    else:
        list(pool.map(one, range(len(lps))))
    rp[M] = Z
    cp[N] = Z
    return dict(m=M, n=N, rp=rp, ci=ci, ar=ar, cp=cp, ri=ri, ac=ac, xr=xr, xc=xc)


def targets(lp: LP, g: dict, x: np.ndarray, y: np.ndarray):
    return x / g["s"], y / g["cs"]


def warm_start(lp: LP, g: dict, xhat: np.ndarray, yhat: np.ndarray):
    """Network outputs -> (x0, y0) for the engine: x clipped to the bounds,
    y projected on the sign each row allows (engine sign: rows without an
    upper bound take y <= 0, rows without a lower bound y >= 0)."""
    x0 = np.clip(np.asarray(xhat, np.float64) * g["s"], lp.var_lower, lp.var_upper)
    y0 = np.asarray(yhat, np.float64) * g["cs"]
    y0 = np.where((lp.row_upper >= INF * 0.1) & (y0 > 0), 0.0, y0)
    y0 = np.where((lp.row_lower <= -INF * 0.1) & (y0 < 0), 0.0, y0)
    return x0, y0
