"""Graph and features of the warm-start network (pm_gnn_lp_features)
against a numpy restatement of their definition (include/pramaan_gnn.h),
on a generated LP, a rolling-horizon window and an LP with free variables,
infinite bounds and an empty row and column."""
import numpy as np
import pytest

INF = 1e30


def features_numpy(lp, period_cols, period_rows):
    m, n = lp.m, lp.n
    b = np.concatenate([lp.row_lower, lp.row_upper, lp.var_lower, lp.var_upper])
    b = np.abs(b[np.abs(b) < 1e29])
    s = max(1.0, float(b.max()) if len(b) else 1.0)
    cs = max(1e-12, float(np.abs(lp.c).max())) if n else 1.0
    rp, ci, v = lp.row_ptr.astype(np.int64), lp.col_idx.astype(np.int64), lp.values
    rowlen = np.diff(rp)
    row_of = np.repeat(np.arange(m), rowlen)
    rmax = np.zeros(m)
    np.maximum.at(rmax, row_of, np.abs(v))
    cmax = np.zeros(n)
    np.maximum.at(cmax, ci, np.abs(v))
    ar = v / np.where(rmax[row_of] > 0, rmax[row_of], 1.0)
    order = np.lexsort((row_of, ci))
    cp = np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n))])
    ri = row_of[order]
    ac = v[order] / np.where(cmax[ci[order]] > 0, cmax[ci[order]], 1.0)
    collen = np.diff(cp)
    fin = lambda a: np.where(np.abs(a) < 1e29, a, 0.0)
    cpos = (np.arange(n) // period_cols) / max(1, n // period_cols - 1) if period_cols else np.zeros(n)
    rpos = (np.arange(m) // period_rows) / max(1, m // period_rows - 1) if period_rows else np.zeros(m)
    def bits(count, period):
        if not period:
            return np.zeros((count, 9))
        role = np.arange(count) % period
        return np.stack([np.where((role >> b) & 1, 1.0, -1.0) for b in range(9)], axis=1)
    xc = np.stack([lp.c / cs, fin(lp.var_lower) / s, fin(lp.var_upper) / s, (lp.var_lower > -1e29) * 1.0,
                   (lp.var_upper < 1e29) * 1.0, collen / max(1, collen.max(initial=0)), cpos], axis=1)
    xr = np.stack([fin(lp.row_lower) / s, fin(lp.row_upper) / s, (lp.row_lower > -1e29) * 1.0,
                   (lp.row_upper < 1e29) * 1.0, (lp.row_lower == lp.row_upper) * 1.0,
                   rowlen / max(1, rowlen.max(initial=0)), rpos], axis=1)
    xc = np.concatenate([xc, bits(n, period_cols)], axis=1)
    xr = np.concatenate([xr, bits(m, period_rows)], axis=1)
    return dict(ar=ar, cp=cp, ri=ri, ac=ac, xr=xr, xc=xc, s=s, cs=cs)


def cases(capi):
    rh = capi.RollingHorizon()
    odd = capi.LP.from_dense([[1.0, 0.0, -2.0], [0.0, 0.0, 0.0], [4.0, 0.0, 0.5]], [3.0, 0.0, -1.0],
                             [-INF, 0.0, 2.0], [5.0, 0.0, 2.0], [-INF, 0.0, -3.0], [INF, 7.0, 1e31])
    return [(capi.generate(None, seed=3, m=120, n=300), 0, 0),
            (rh.window(7), rh.cols_per_period, rh.rows_per_period),
            (odd, 0, 0)]


def test_features_match_definition(capi):
    from pramaan_amd import warmstart as W
    for lp, pc, pr in cases(capi):
        g = W.graph_of(lp, pc, pr)
        want = features_numpy(lp, pc, pr)
        assert g["s"] == want["s"] and g["cs"] == want["cs"]
        assert np.array_equal(g["cp"], want["cp"]) and np.array_equal(g["ri"], want["ri"])
        for k in ("ar", "ac", "xr", "xc"):
            assert np.array_equal(g[k], want[k].astype(np.float32)), k


def test_features_reject_bad_lp(capi):
    from pramaan_amd import warmstart as W
    lp = capi.generate(None, seed=3, m=20, n=40)
    lp.col_idx[3] = 40
    with pytest.raises(capi.PramaanError):
        W.graph_of(lp)


def test_batch_graphs_in_place_match_concatenation(capi):
    """batch_graphs_of (each LP's graph written straight into its slices of
    the combined arrays, in parallel) gives the arrays of batch_graphs over
    graph_of, an LP without rows included."""
    from concurrent.futures import ThreadPoolExecutor
    from pramaan_amd import warmstart as W
    rh = capi.RollingHorizon()
    lps = [rh.window(w) for w in (0, 1, 250)]
    lps.insert(1, capi.LP.from_dense(np.zeros((0, 3)), [1.0, -1.0, 2.0], [], [], np.zeros(3), np.ones(3)))
    want = W.batch_graphs([W.graph_of(lp, rh.cols_per_period, rh.rows_per_period) for lp in lps])
    with ThreadPoolExecutor(4) as pool:
        got = W.batch_graphs_of(lps, rh.cols_per_period, rh.rows_per_period, pool)
    assert (got["m"], got["n"]) == (want["m"], want["n"])
    for k in ("rp", "ci", "ar", "cp", "ri", "ac", "xr", "xc"):
        assert np.array_equal(got[k], want[k]), k
