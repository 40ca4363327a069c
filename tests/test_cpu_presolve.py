"""Ruiz equilibration on the CPU: the oracle's restatement of the reference's
ruiz_scale (src/presolve/scaling.cpp:29-150), and the argument checks of
pm_ruiz_scale (which return before any device call).

The oracle is pinned two ways:
  - against the reference's own ruiz_scale, compiled from its sources into
    oracle/_ref/ref_driver, on afiro and adlittle: every factor, scaled
    entry, cost and bound equal to the last bit;
  - by what the scaling must satisfy whatever the code: the scaled entries
    are A[i, j] * row_scale[i] * col_scale[j] (a transposed or dropped
    factor fails), the finite bounds move by the row factor and the inverse
    column factor while absent ones (|bound| >= 1e30) stay, and many passes
    drive every row norm and every column norm (including |c_j|) to 1.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
REF = ROOT / "oracle" / "_ref"
INF = 1e30


def read(name):
    from pramaan_amd import mps
    return mps.read_mps(GOLDEN / name)


def reference_ruiz(name, passes, tmp_path):
    """Output of the reference's ruiz_scale on tests/golden/<name>."""
    out = tmp_path / "ruiz.txt"
    r = subprocess.run([str(REF / "ref_driver"), "ruiz", str(GOLDEN / name), str(passes), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out.read_text().split("\n")


@pytest.mark.skipif(not (REF / "ref_driver").exists(), reason="reference CPU programs (oracle/_ref) not built")
@pytest.mark.parametrize("name", ["afiro.mps", "adlittle.mps"])
@pytest.mark.parametrize("passes", [1, 10])
def test_oracle_ruiz_equals_reference_build(oracle, name, passes, tmp_path):
    lp = read(name)
    lines = reference_ruiz(name, passes, tmp_path)
    o = oracle.ruiz_scale(lp, passes)
    m, n = lp.m, lp.n
    rs = np.array([float(x) for x in lines[:m]])
    cs = np.array([float(x) for x in lines[m:m + n]])
    nnz = len(lp.values)
    ent = [ln.split() for ln in lines[m + n:m + n + nnz]]
    rest = lines[m + n + nnz:]
    c = np.array([float(x) for x in rest[:n]])
    rb = np.array([[float(t) for t in ln.split()] for ln in rest[n:n + m]])
    vb = np.array([[float(t) for t in ln.split()] for ln in rest[n + m:n + m + n]])
    assert np.array_equal(o["row_scale"], rs) and np.array_equal(o["col_scale"], cs)
    assert np.array_equal(o["c"], c)
    assert np.array_equal(o["row_lower"], rb[:, 0]) and np.array_equal(o["row_upper"], rb[:, 1])
    assert np.array_equal(o["var_lower"], vb[:, 0]) and np.array_equal(o["var_upper"], vb[:, 1])
    # entries row by row as (column, value) sets: the two readers may order
    # a row's entries differently
    k = 0
    for i in range(m):
        lo, hi = lp.row_ptr[i], lp.row_ptr[i + 1]
        mine = sorted(zip(lp.col_idx[lo:hi].tolist(), o["values"][lo:hi].tolist()))
        theirs = sorted((int(e[0]), float(e[1])) for e in ent[k:k + hi - lo])
        assert mine == theirs, i
        k += hi - lo


def _dense_lp(capi, A, c, rl, ru, vl, vu):
    return capi.LP.from_dense(np.asarray(A, float), c, rl, ru, vl, vu)


def test_oracle_ruiz_invariants(capi, oracle):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 60)) * (rng.random((40, 60)) < 0.2) * 10.0 ** rng.integers(-4, 5, (40, 60))
    A[3] = 0.0  # an empty row
    A[:, 7] = 0.0  # an empty column
    c = rng.standard_normal(60) * 10.0 ** rng.integers(-3, 4, 60)
    rl = np.where(rng.random(40) < 0.3, -INF, rng.standard_normal(40))
    ru = np.where(rng.random(40) < 0.3, INF, rl + 1.0)
    ru[ru < rl] = INF
    vl = np.where(rng.random(60) < 0.5, 0.0, -INF)
    vu = np.where(rng.random(60) < 0.5, 5.0, 2e30)
    lp = _dense_lp(capi, A, c, rl, ru, vl, vu)
    o = oracle.ruiz_scale(lp, 60)
    rs, cs = o["row_scale"], o["col_scale"]
    rows = np.repeat(np.arange(lp.m), np.diff(lp.row_ptr))
    want = lp.values * rs[rows] * cs[lp.col_idx]
    assert np.allclose(o["values"], want, rtol=1e-12, atol=0)
    assert np.allclose(o["c"], lp.c * cs, rtol=1e-12, atol=0)
    fin = np.abs(lp.row_lower) < INF
    assert np.allclose(o["row_lower"][fin], lp.row_lower[fin] * rs[fin], rtol=1e-12)
    assert np.array_equal(o["row_lower"][~fin], lp.row_lower[~fin])
    fin = np.abs(lp.var_upper) < INF
    assert np.allclose(o["var_upper"][fin], lp.var_upper[fin] / cs[fin], rtol=1e-12)
    assert np.array_equal(o["var_upper"][~fin], lp.var_upper[~fin])
    assert rs[3] == 1.0  # empty row: norm 0, factor 1 every pass
    S = np.zeros((lp.m, lp.n))
    S[rows, lp.col_idx] = np.abs(o["values"])
    rn = S.max(axis=1)
    cn = np.maximum(S.max(axis=0), np.abs(o["c"]))
    assert np.allclose(rn[rn > 0], 1.0, atol=1e-6)
    assert np.allclose(cn[cn > 0], 1.0, atol=1e-6)


def test_ruiz_rejects_bad_input_and_empty_models(capi):
    """Argument errors are reported before any device call; an LP without
    rows or columns is left as it is with unit factors (scaling.cpp:37)."""
    import ctypes as C
    lib = capi.lib()
    lp = _dense_lp(capi, [[1.0, 2.0]], [1.0, 1.0], [0.0], [1.0], [0.0, 0.0], [1.0, 1.0])
    bad = lp.col_idx.copy()
    bad[1] = 5
    d = lambda a: np.ascontiguousarray(a, np.float64)
    v, c, rl, ru, vl, vu = (d(lp.values), d(lp.c), d(lp.row_lower), d(lp.row_upper), d(lp.var_lower),
                            d(lp.var_upper))
    rs, cs = np.zeros(1), np.zeros(2)
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    rp = np.ascontiguousarray(lp.row_ptr, np.int32)
    rc = lib.pm_ruiz_scale(1, 2, p(rp, C.c_int32), p(np.ascontiguousarray(bad, np.int32), C.c_int32),
                           p(v, C.c_double), p(c, C.c_double), p(rl, C.c_double), p(ru, C.c_double),
                           p(vl, C.c_double), p(vu, C.c_double), 10, p(rs, C.c_double), p(cs, C.c_double))
    assert rc == 1 and "column index out of range" in lib.pm_last_error().decode()
    rc = lib.pm_ruiz_scale(1, 2, p(rp, C.c_int32), p(np.ascontiguousarray(lp.col_idx, np.int32), C.c_int32),
                           p(v, C.c_double), p(c, C.c_double), p(rl, C.c_double), p(ru, C.c_double),
                           p(vl, C.c_double), p(vu, C.c_double), -1, p(rs, C.c_double), p(cs, C.c_double))
    assert rc == 1
    empty = _dense_lp(capi, np.zeros((0, 3)), [1.0, -2.0, 3.0], [], [], np.zeros(3), np.ones(3))
    r = capi.ruiz_scale(empty)
    assert np.array_equal(r["col_scale"], np.ones(3)) and np.array_equal(r["c"], empty.c)
    assert np.array_equal(r["var_upper"], np.ones(3))
