"""Learned warm start for rolling-horizon LPs: graph features, targets, and
the mapping between network outputs and PDHG starting points.

The network itself (pm_gnn_*, HIP) sees one LP as a bipartite graph:
  column features (8): c_j / max|c|, l_j / s, u_j / s (0 when infinite),
                       1[l finite], 1[u finite], nnz_j / max nnz,
                       position of the column's period in the window, 1
  row features (8):    l_i / s, u_i / s (0 when infinite), 1[l finite],
                       1[u finite], 1[equality], nnz_i / max nnz,
                       period position, 1
with s = max(1, largest finite |bound|).  Targets are x_j / s and
y_i / max|c| (y in the engine's sign).  Several LPs are batched as one
block-diagonal graph.
"""
from __future__ import annotations

import numpy as np

from ._capi import INF, LP

FEAT = 8


def scales(lp: LP) -> tuple[float, float]:
    b = np.concatenate([lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper])
    b = np.abs(b[np.abs(b) < 1e29])
    s = max(1.0, float(b.max()) if len(b) else 1.0)
    cs = max(1e-12, float(np.abs(lp.c).max()) if lp.n else 1.0)
    return s, cs


def graph_of(lp: LP, period_cols: int | None = None, period_rows: int | None = None) -> dict:
    """Graph arrays of one LP (see the module docstring)."""
    m, n = lp.m, lp.n
    s, cs = scales(lp)
    rp, ci, v = lp.row_ptr.astype(np.int64), lp.col_idx.astype(np.int64), lp.values
    rowlen = np.diff(rp)
    row_of = np.repeat(np.arange(m), rowlen)
    absv = np.abs(v)
    rmax = np.zeros(m)
    np.maximum.at(rmax, row_of, absv)
    cmax = np.zeros(n)
    np.maximum.at(cmax, ci, absv)
    ar = v / np.where(rmax[row_of] > 0, rmax[row_of], 1.0)
    order = np.lexsort((row_of, ci))  # column-major, rows ascending
    cp = np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n))])
    ri = row_of[order]
    ac = v[order] / np.where(cmax[ci[order]] > 0, cmax[ci[order]], 1.0)
    collen = np.diff(cp)
    fin = lambda a: np.where(np.abs(a) < 1e29, a, 0.0)
    cpos = (np.arange(n) // period_cols) / max(1, (n // period_cols) - 1) if period_cols else np.zeros(n)
    rpos = (np.arange(m) // period_rows) / max(1, (m // period_rows) - 1) if period_rows else np.zeros(m)
    xc = np.stack([lp.c / cs, fin(lp.var_lower) / s, fin(lp.var_upper) / s, (lp.var_lower > -1e29) * 1.0,
                   (lp.var_upper < 1e29) * 1.0, collen / max(1, collen.max()), cpos, np.ones(n)], axis=1)
    xr = np.stack([fin(lp.row_lower) / s, fin(lp.row_upper) / s, (lp.row_lower > -1e29) * 1.0,
                   (lp.row_upper < 1e29) * 1.0, (lp.row_lower == lp.row_upper) * 1.0,
                   rowlen / max(1, rowlen.max()), rpos, np.ones(m)], axis=1)
    return dict(m=m, n=n, rp=rp, ci=ci, ar=ar, cp=cp, ri=ri, ac=ac, xr=xr, xc=xc, s=s, cs=cs)


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
