"""GPU kernel parity: SpMV, transpose SpMV and the Step-1 PDHG update,
HIP path (through the C-ABI) against the CPU oracle.

Cases restate the reference's tests/test_gpu_pdhg.cpp:236-258 (diagonal
d x d for d in {1,2,31,...,129}, the 3x7 rectangle, m = 0) and
tests/test_pdhg_kernel.cpp:129-241 (edge cases, dimensions 0..1025,
mt19937-style randomised trials, numerical stress, repeated calls).
Tolerances: the reference's 1e-4 + 1e-4 * max(|a|,|b|) for SpMV
(test_gpu_pdhg.cpp:192,224) and 1e-5 + 1e-4 rel for the updates
(test_pdhg_kernel.cpp:125-126); the sequential variant must match the
oracle bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [1, 2, 31, 32, 33, 63, 64, 65, 127, 128, 129]
VARIANTS = ["sequential", "scalar", "vector2", "vector4", "vector8", "vector16", "vector32", "vector64", "merge",
            "auto"]


def dense_to_csr(A):
    A = np.asarray(A, dtype=np.float64)
    rp, ci, v = [0], [], []
    for row in A:
        for j, a in enumerate(row):
            if a != 0:
                ci.append(j)
                v.append(a)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v, np.float32)


def close(ref, got, atol=1e-4, rtol=1e-4):
    ref = np.asarray(ref, np.float64)
    got = np.asarray(got, np.float64)
    tol = atol + rtol * np.maximum(np.abs(ref), np.abs(got))
    return bool(np.all(np.abs(ref - got) <= tol))


def cases():
    out = []
    for d in DIMS:
        out.append((f"diag{d}", d, d, np.diag(np.arange(1, d + 1, dtype=float))))
    R = np.zeros((3, 7))
    R[0, 1], R[0, 5], R[1, 3], R[2, 0], R[2, 6] = 2.0, -1.0, 3.0, 1.0, 4.0
    out.append(("rect3x7", 3, 7, R))
    return out


@pytest.mark.parametrize("variant", VARIANTS)
def test_spmv_reference_shapes(capi, oracle, variant):
    for name, m, n, A in cases():
        rp, ci, v = dense_to_csr(A)
        x = (np.arange(n, dtype=np.float32) + 1) * 0.5
        ref = oracle.spmv(rp, ci, v, x, m)
        got = capi.launch_spmv(rp, ci, v, x, m, n, variant)
        assert got.shape == (m,), name
        if variant == "sequential":
            assert np.array_equal(ref, got), name
        assert close(ref, got), name
        xt = (np.arange(m, dtype=np.float32) + 1) * 0.3
        ref_t = oracle.spmv_t(rp, ci, v, xt, m, n)
        got_t = capi.launch_spmv_t(rp, ci, v, xt, m, n, variant)
        assert got_t.shape == (n,), name
        if variant == "sequential":
            assert np.array_equal(ref_t, got_t), name
        assert close(ref_t, got_t), name


def test_spmv_empty_rows(capi):
    rp = np.zeros(1, np.int32)
    y = capi.launch_spmv(rp, np.zeros(0, np.int32), np.zeros(0, np.float32), np.ones(5, np.float32), 0, 5)
    assert y.shape == (0,)
    yt = capi.launch_spmv_t(rp, np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros(0, np.float32), 0, 5)
    assert yt.shape == (5,) and np.all(yt == 0)


@pytest.mark.parametrize("variant", VARIANTS)
def test_spmv_synthetic_lp(capi, oracle, variant):
    lp = capi.generate(0, seed=7)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(lp.n).astype(np.float32)
    y = rng.standard_normal(lp.m).astype(np.float32)
    ref = oracle.spmv(lp.row_ptr, lp.col_idx, lp.values, x, lp.m)
    got = capi.launch_spmv(lp.row_ptr, lp.col_idx, lp.values, x, lp.m, lp.n, variant)
    ref_t = oracle.spmv_t(lp.row_ptr, lp.col_idx, lp.values, y, lp.m, lp.n)
    got_t = capi.launch_spmv_t(lp.row_ptr, lp.col_idx, lp.values, y, lp.m, lp.n, variant)
    if variant == "sequential":
        assert np.array_equal(ref, got) and np.array_equal(ref_t, got_t)
    assert close(ref, got) and close(ref_t, got_t)


def test_spmv_ragged_rows(capi, oracle):
    """Rows of length 0..300 (one long row, many empty ones)."""
    rng = np.random.default_rng(11)
    m, n = 700, 400
    lens = rng.integers(0, 12, m)
    lens[::97] = 0
    lens[5] = 300
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    v = rng.standard_normal(len(ci)).astype(np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    ref = oracle.spmv(rp, ci, v, x, m)
    for variant in VARIANTS:
        assert close(ref, capi.launch_spmv(rp, ci, v, x, m, n, variant)), variant


def test_spmv_merge_skewed_rows(capi, oracle):
    """Merge-path SpMV on Zipf row lengths (3.8M nonzeros over 20,000 rows:
    median 1, 44 % empty, 63 rows of 10^4 to 10^5 entries), where a row
    spans many threads' ranges and many rows fit in one: against the
    oracle, in both orientations."""
    rng = np.random.default_rng(5)
    m, n = 20000, 30000
    lens = np.minimum(rng.zipf(1.6, m) - 1, 100000)
    lens[[3, 777, 19999]] = [100000, 54321, 12345]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = rng.integers(0, n, int(rp[-1])).astype(np.int32)
    v = rng.standard_normal(len(ci)).astype(np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(m).astype(np.float32)
    ref = oracle.spmv(rp, ci, v, x, m)
    got = capi.launch_spmv(rp, ci, v, x, m, n, "merge")
    # rows of up to 10^5 products: FP32 sums in a different order than the
    # oracle's, so the bound scales with each row's sum of |a_ij x_j|
    mag = oracle.spmv(rp, ci, np.abs(v), np.abs(x), m)
    assert np.all(np.abs(np.asarray(got, np.float64) - ref) <= 1e-4 + 1e-5 * np.asarray(mag, np.float64))
    ref_t = oracle.spmv_t(rp, ci, v, y, m, n)
    mag_t = oracle.spmv_t(rp, ci, np.abs(v), np.abs(y), m, n)
    got_t = capi.launch_spmv_t(rp, ci, v, y, m, n, "merge")
    assert np.all(np.abs(np.asarray(got_t, np.float64) - ref_t) <= 1e-4 + 1e-5 * np.asarray(mag_t, np.float64))


def step1_case(capi, oracle, x, y, g, axb, b, tau, sigma):
    ex, ey = oracle.step1_iteration(x, y, g, axb, b, tau, sigma)
    gx, gy = capi.launch_pdhg_iteration(x, y, g, axb, b, tau, sigma)
    assert np.array_equal(ex, gx) and np.array_equal(ey, gy)
    assert close(ex, gx, 1e-5, 1e-4) and close(ey, gy, 1e-5, 1e-4)


def test_step1_edge_cases(capi, oracle):
    x = [1.0, 1.0, 2.0, 1.0, 0.0, 5.0, 1.0, 1.0]
    g = [0.5, -0.5, 2.0, 3.0, 1.0, 0.0, 1e5, -1e5]
    step1_case(capi, oracle, x, [], g, [], [], 1.0, 1.0)
    y = [0.0, 0.0, 0.0, 5.0, 0.0, 2.0, 1e5, 1e-5, -2.0]
    axb = [1.0, -1.0, 1.0, 2.0, 1.0, 0.0, -1e5, 1e-5, 0.0]
    b = [0.0, 0.0, 1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    step1_case(capi, oracle, [], y, [], axb, b, 1.0, 1.0)
    step1_case(capi, oracle, [], y, [], axb, b, 1.0, 0.0)


def test_step1_dimensions_and_empty(capi, oracle):
    for d in [0, 1, 2, 3, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1024, 1025]:
        step1_case(capi, oracle, np.ones(d), np.ones(d), np.full(d, 0.5), np.full(d, 0.5), np.zeros(d), 1.0, 1.0)
    step1_case(capi, oracle, [1.0], [1.0], [1.0], [1.0], [1.0], 1.0, 1.0)
    step1_case(capi, oracle, [], [1.0], [], [1.0], [1.0], 1.0, 1.0)
    step1_case(capi, oracle, [1.0], [], [1.0], [], [], 1.0, 1.0)
    step1_case(capi, oracle, [], [], [], [], [], 1.0, 1.0)


def test_step1_randomised_and_stress(capi, oracle):
    rng = np.random.default_rng(42)
    for _ in range(50):
        x = rng.uniform(-10, 10, 150).astype(np.float32)
        x[::7] = 0.0
        g = rng.uniform(-10, 10, 150).astype(np.float32)
        y, axb, b = (rng.uniform(-10, 10, 120).astype(np.float32) for _ in range(3))
        b[::7] = 0.0
        tau, sigma = abs(rng.uniform(-10, 10)), abs(rng.uniform(-10, 10))
        step1_case(capi, oracle, x, y, g, axb, b, tau, sigma)
    step1_case(capi, oracle, [1e-10, 1e10, 1e-38, 1.0, 0.0, -0.0], [1e-10, 1e10, 1e-38, -1e-38],
               [1e-10, -1e10, -1e-38, 1.0, 1e-38, -1e-38], [1e-10, 1e10, 1e-38, -1e-38],
               [1e-10, 1e10, 1e-38, -1e-38], 1.0, 1.0)


def test_step1_repeated(capi, oracle):
    x = np.ones(1000, np.float32)
    y = np.ones(1000, np.float32)
    g = np.full(1000, 0.1, np.float32)
    a = np.full(1000, 0.1, np.float32)
    b = np.zeros(1000, np.float32)
    cx, cy = x.copy(), y.copy()
    for _ in range(50):
        x, y = capi.launch_pdhg_iteration(x, y, g, a, b, 1.0, 1.0)
        cx, cy = oracle.step1_iteration(cx, cy, g, a, b, 1.0, 1.0)
    assert np.array_equal(x, cx) and np.array_equal(y, cy)


def test_invalid_csr_is_rejected(capi):
    rp = np.array([0, 2, 1], np.int32)
    with pytest.raises(capi.PramaanError):
        capi.launch_spmv(rp, np.array([0], np.int32), np.ones(1, np.float32), np.ones(2, np.float32), 2, 2)
