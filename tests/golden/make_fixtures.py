"""Writes tests/golden/fixtures.json: FP64 optimal objectives of small LPs from
scipy's HiGHS (linprog, method="highs"), computed in the build container
only.  Covers rolling-horizon blending windows (default plan and a perturbed
one), whose optimum has no closed form, and the Netlib afiro / adlittle
files.  Test-only data; the engine never calls HiGHS.

Usage: python tests/golden/make_fixtures.py
"""
import json
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from pramaan_amd import _capi as P  # noqa: E402
from pramaan_amd.mps import read_mps  # noqa: E402

INF = 1e29


def highs(lp):
    A = sp.csr_matrix((lp.values, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))
    eq = lp.row_lower == lp.row_upper
    hi = (~eq) & (lp.row_upper < INF)
    lo = (~eq) & (lp.row_lower > -INF)
    A_ub = sp.vstack([A[hi], -A[lo]]).tocsr()
    b_ub = np.concatenate([lp.row_upper[hi], -lp.row_lower[lo]])
    c = -lp.c if lp.sense == 1 else lp.c
    bounds = [(None if l <= -INF else l, None if u >= INF else u) for l, u in zip(lp.var_lower, lp.var_upper)]
    r = linprog(c, A_ub=A_ub if A_ub.shape[0] else None, b_ub=b_ub if A_ub.shape[0] else None,
                A_eq=A[eq] if eq.any() else None, b_eq=lp.row_lower[eq] if eq.any() else None,
                bounds=bounds, method="highs")
    assert r.status == 0, r.message
    obj = (-r.fun if lp.sense == 1 else r.fun) + lp.offset
    return float(obj)


def main():
    out = {"source": "scipy.optimize.linprog(method='highs')", "rolling": {}, "rolling_perturbed": {}, "netlib": {}}
    rh = P.RollingHorizon()
    for w in (0, 1, 2, 100, 250, 499):
        out["rolling"][str(w)] = highs(rh.window(w))
    rp = P.RollingHorizon(seed=7, demand_noise=0.2, blend_noise=0.1, perturb_seed=11)
    for w in (0, 37, 400):
        out["rolling_perturbed"][str(w)] = highs(rp.window(w))
    for name in ("afiro", "adlittle"):
        out["netlib"][name] = highs(read_mps(HERE / f"{name}.mps"))
    (HERE / "fixtures.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
