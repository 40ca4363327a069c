"""write_mps (pramaan_amd/mps.py) against read_mps, the reader that follows
the reference's parse_mps: a written LP reads back with the same rows,
columns, values, bounds and objective (a maximisation as the minimisation
of -c), and rows with two different finite bounds are refused (the format
would need RANGES, which the reference's parser rejects)."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def same(a, b, sign=1.0):
    assert (a.m, a.n, a.nnz) == (b.m, b.n, b.nnz)
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col_idx, b.col_idx)
    assert np.array_equal(a.values, b.values)
    assert np.array_equal(sign * a.c, b.c)
    for f in ("row_lower", "row_upper", "var_lower", "var_upper"):
        x, y = getattr(a, f), getattr(b, f)
        assert np.array_equal(np.where(np.abs(x) >= 1e29, np.sign(x), x), np.where(np.abs(y) >= 1e29, np.sign(y), y)), f


@pytest.mark.parametrize("name", ["afiro", "adlittle"])
def test_netlib_round_trip(capi, tmp_path, name):
    from pramaan_amd.mps import read_mps, write_mps
    lp = read_mps(GOLDEN / f"{name}.mps")
    write_mps(lp, tmp_path / "a.mps")
    same(lp, read_mps(tmp_path / "a.mps"))
    assert read_mps(tmp_path / "a.mps").offset == lp.offset


def test_bounds_and_sense(capi, tmp_path):
    from pramaan_amd.mps import read_mps, write_mps
    A = [[1.0, 2.0, 0.0, 1.0], [0.0, 1.0, -1.0, 0.0], [3.0, 0.0, 0.0, 1.0], [1.0, 1.0, 1.0, 1.0]]
    inf = 1e30
    lp = capi.LP.from_dense(A, [1.0, -2.0, 0.5, 0.0], [-inf, 1.0, 2.0, -inf], [4.0, 1.0, inf, inf],
                            [0.0, -inf, -3.0, 2.0], [5.0, inf, inf, 2.0])
    write_mps(lp, tmp_path / "b.mps")
    same(lp, read_mps(tmp_path / "b.mps"))
    lp.sense = 1
    write_mps(lp, tmp_path / "c.mps")
    same(lp, read_mps(tmp_path / "c.mps"), sign=-1.0)


def test_ranged_row_is_refused(capi, tmp_path):
    from pramaan_amd.mps import MpsError, write_mps
    lp = capi.LP.from_dense([[1.0, 1.0]], [1.0, 1.0], [1.0], [2.0], [0.0, 0.0], [1e30, 1e30])
    with pytest.raises(MpsError):
        write_mps(lp, tmp_path / "d.mps")


def test_rolling_window_round_trip(capi, tmp_path):
    from pramaan_amd.mps import read_mps, write_mps
    lp = capi.RollingHorizon().window(7)
    write_mps(lp, tmp_path / "w.mps")
    same(lp, read_mps(tmp_path / "w.mps"))
