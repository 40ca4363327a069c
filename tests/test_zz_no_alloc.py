"""pm_ruiz_scale_noalloc (device arrays in, no allocation; scratch in a
fixed device array of the library) against pm_ruiz_scale on the same LPs:
every scaled value, cost, bound and cumulative factor bit for bit, on two
calls in a row with the second LP smaller than the first (the shared
scratch keeps the first call's contents), plus the size limit and the
guarded rejection of a bad column index."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("values", "c", "row_lower", "row_upper", "var_lower", "var_upper", "row_scale", "col_scale")
H2D, D2H = 1, 2  # hipMemcpyKind


class DeviceArrays:
    """The caller's device arrays, allocated through the HIP runtime the
    engine links (a second runtime in the process, such as the one torch
    bundles, would not see the device once this one holds it)."""

    def __init__(self, lp):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.host = {k: np.ascontiguousarray(getattr(lp, k)) for k in
                     ("row_ptr", "col_idx", "values", "c", "row_lower", "row_upper", "var_lower", "var_upper")}
        self.host["row_scale"] = np.zeros(max(lp.m, 1))
        self.host["col_scale"] = np.zeros(max(lp.n, 1))
        self.ptr = {}
        for k, a in self.host.items():
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
            self.ptr[k] = p.value
            self.put(k)

    def put(self, k):
        a = self.host[k]
        assert self.hip.hipMemcpy(C.c_void_p(self.ptr[k]), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                  H2D) == 0

    def get(self, k):
        a = np.empty_like(self.host[k])
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr[k]), C.c_size_t(a.nbytes),
                                  D2H) == 0
        return a

    def free(self):
        for p in self.ptr.values():
            self.hip.hipFree(C.c_void_p(p))
        self.ptr = {}


def run_noalloc(capi, lp, passes=10):
    d = DeviceArrays(lp)
    try:
        capi.ruiz_scale_noalloc(lp.m, lp.n, lp.nnz, d.ptr, passes)
        out = {k: d.get(k) for k in KEYS}
    finally:
        d.free()
    out["row_scale"], out["col_scale"] = out["row_scale"][:lp.m], out["col_scale"][:lp.n]
    return out


def test_noalloc_equals_allocating_entry_on_two_calls(capi):
    big = capi.generate(2, seed=1)          # 50,000 x 100,000, 1.2M nonzeros
    small = capi.generate(0, seed=2)        # 2,000 x 5,000
    assert small.m < big.m and small.n < big.n
    for lp in (big, small):
        print(f"noalloc ruiz m={lp.m} n={lp.n} nnz={lp.nnz}", flush=True)
        got = run_noalloc(capi, lp)
        want = capi.ruiz_scale(lp, 10)
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (lp.m, k)


def test_noalloc_rejects_oversize_and_bad_index(capi):
    lp = capi.generate(None, seed=5, m=40, n=90)
    d = DeviceArrays(lp)
    try:
        with pytest.raises(capi.PramaanError) as e:     # checked on the host: no launch
            capi.ruiz_scale_noalloc(1, 2_100_000, lp.nnz, d.ptr)
        assert "exceeds the scratch" in str(e.value)
        d.host["col_idx"] = d.host["col_idx"].copy()
        d.host["col_idx"][7] = lp.n + 3                 # out of range: flagged, never dereferenced
        d.put("col_idx")
        with pytest.raises(capi.PramaanError) as e:
            capi.ruiz_scale_noalloc(lp.m, lp.n, lp.nnz, d.ptr)
        assert "outside the sizes given" in str(e.value)
    finally:
        d.free()
    # the next valid call on the same scratch is unaffected
    got = run_noalloc(capi, lp)
    want = capi.ruiz_scale(lp, 10)
    assert all(np.array_equal(got[k], want[k]) for k in KEYS)
