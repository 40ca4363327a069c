import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP path through the C-ABI")
    config.addinivalue_line("markers", "slow: larger instances")


@pytest.fixture(scope="session")
def capi():
    from pramaan_amd import _capi
    _capi.lib()
    return _capi


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.lib()
    return binding


def _ruiz_pc_restated(lp, passes=10):
    """Ruiz infinity-norm passes on A, then Pock-Chambolle 1-norms, as plain
    loops in CSR order (rows) and in row order per column -- the order the
    device scaling of both engines follows."""
    m, n = lp.m, lp.n
    rp, ci = lp.row_ptr.tolist(), lp.col_idx.tolist()
    v = lp.values.astype(float).tolist()
    dr, dc = [1.0] * m, [1.0] * n
    for p in range(passes + 1):
        rn, cn = [0.0] * m, [0.0] * n
        for i in range(m):
            for k in range(rp[i], rp[i + 1]):
                a = abs(v[k])
                if p < passes:
                    rn[i] = max(rn[i], a)
                    cn[ci[k]] = max(cn[ci[k]], a)
                else:
                    rn[i] += a
                    cn[ci[k]] += a
        rn = [1.0 / np.sqrt(a) if a > 0 else 1.0 for a in rn]
        cn = [1.0 / np.sqrt(a) if a > 0 else 1.0 for a in cn]
        for i in range(m):
            for k in range(rp[i], rp[i + 1]):
                v[k] *= rn[i] * cn[ci[k]]
        dr = [d * r for d, r in zip(dr, rn)]
        dc = [d * c for d, c in zip(dc, cn)]
    return np.array(dr), np.array(dc)


@pytest.fixture(scope="session")
def ruiz_restated():
    return _ruiz_pc_restated
