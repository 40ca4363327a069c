// launchers.hip -- test-facing entry points that take host vectors, run one
// kernel and copy the result back: the reference's launch_spmv,
// launch_spmv_t and launch_pdhg_iteration (src/gpu/pdhg.cu:77-241).  They
// reuse the solver's row kernels, so the SpMV tests exercise the same code
// as the hot loop.
#include <vector>

#include "dispatch.hpp"

namespace pm {
namespace {

// Step-1 primal update (pdhg.cu:40-54): x = max(0, x - tau g).
struct FnStep1Primal {
    float* x;
    const float* g;
    float tau;
    __device__ void operator()(int j) const { x[j] = fmaxf(0.0f, x[j] - tau * g[j]); }
};
// Step-1 dual update (pdhg.cu:56-66): y = y + sigma (Ax_bar - b).
struct FnStep1Dual {
    float* y;
    const float* axb;
    const float* b;
    float sigma;
    __device__ void operator()(int i) const { y[i] = y[i] + sigma * (axb[i] - b[i]); }
};

int pick_variant(int variant, int nrows, int nnz) {
    if (variant < 0 || variant >= PM_SPMV_NUM_VARIANTS) fail(PM_ERR_INVALID_ARG, "unknown SpMV variant");
    if (variant != PM_SPMV_AUTO) return variant;
    const double avg = nrows > 0 ? static_cast<double>(nnz) / nrows : 0.0;
    if (avg <= 3.0) return PM_SPMV_SCALAR;
    if (avg <= 8.0) return PM_SPMV_VECTOR2;
    if (avg <= 16.0) return PM_SPMV_VECTOR4;
    if (avg <= 40.0) return PM_SPMV_VECTOR8;
    if (avg <= 100.0) return PM_SPMV_VECTOR16;
    return PM_SPMV_VECTOR32;
}

void check_csr(const int32_t* rp, const int32_t* ci, const float* v, int32_t nnz, int32_t m, int32_t n) {
    if (m < 0 || n < 0 || nnz < 0) fail(PM_ERR_INVALID_ARG, "negative dimension");
    if (!rp) fail(PM_ERR_INVALID_ARG, "row_ptr is NULL");
    if (nnz > 0 && (!ci || !v)) fail(PM_ERR_INVALID_ARG, "col_idx / values are NULL");
    if (rp[0] != 0 || rp[m] != nnz) fail(PM_ERR_INVALID_ARG, "row_ptr[0] must be 0 and row_ptr[m] == nnz");
    for (int i = 0; i < m; ++i)
        if (rp[i] > rp[i + 1]) fail(PM_ERR_INVALID_ARG, "row_ptr must be non-decreasing");
    for (int k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= n) fail(PM_ERR_INVALID_ARG, "col_idx entry out of bounds");
}

// Runs y[rows] = M x over a host CSR (rows x cols) on the null stream.
void run_spmv(const int32_t* rp, const int32_t* ci, const float* v, int32_t nnz, const float* x, float* y,
              int32_t rows, int32_t cols, int32_t variant) {
    const int var = pick_variant(variant, rows, nnz);  // checked before the device is touched
    if (rows == 0) return;
    PM_HIP(hipSetDevice(0));
    hipStream_t s = nullptr;
    DevBuf<int> d_rp, d_ci;
    DevBuf<float> d_v, d_x, d_y(rows);
    d_rp.upload(rp, static_cast<size_t>(rows) + 1, s);
    d_ci.upload(ci, nnz, s);
    d_v.upload(v, nnz, s);
    d_x.upload(x, cols, s);
    if (cols == 0) d_x.resize(1);
    if (var == PM_SPMV_MERGE) {
        const size_t nt = static_cast<size_t>(merge_tiles(rows, nnz));
        DevBuf<int> prow(nt + 1);
        DevBuf<float> carry(nt), head(nt);
        const std::vector<int4> longs = merge_long_rows(rows, rp);
        DevBuf<int4> d_longs;
        d_longs.upload(longs, s);
        launch_merge_partition(rows, nnz, d_rp.get(), prow.get(), s);
        launch_merge(rows, nnz, d_rp.get(), d_ci.get(), d_v.get(), d_x.get(), EpiStore<float>{d_y.get()}, prow.get(),
                     d_longs.get(), static_cast<int>(longs.size()), carry.get(), head.get(), s);
        d_y.download(y, rows, s);
        PM_HIP(hipStreamSynchronize(s));
        return;
    }
    launch_rows<float>(var, rows, d_rp.get(), d_ci.get(), d_v.get(), d_x.get(), EpiStore<float>{d_y.get()}, s);
    d_y.download(y, rows, s);
    PM_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace pm

extern "C" {

int pm_launch_spmv(const int32_t* row_ptr, const int32_t* col_idx, const float* values, int32_t nnz,
                   const float* x, float* y, int32_t m, int32_t n, int32_t variant) {
    return pm::guarded([&] {
        pm::check_csr(row_ptr, col_idx, values, nnz, m, n);
        if (m > 0 && !y) pm::fail(PM_ERR_INVALID_ARG, "y is NULL");
        if (n > 0 && !x) pm::fail(PM_ERR_INVALID_ARG, "x is NULL");
        pm::run_spmv(row_ptr, col_idx, values, nnz, x, y, m, n, variant);
    });
}

int pm_launch_spmv_t(const int32_t* row_ptr, const int32_t* col_idx, const float* values, int32_t nnz,
                     const float* x, float* y, int32_t m, int32_t n, int32_t variant) {
    return pm::guarded([&] {
        pm::check_csr(row_ptr, col_idx, values, nnz, m, n);
        if (n > 0 && !y) pm::fail(PM_ERR_INVALID_ARG, "y is NULL");
        if (m > 0 && !x) pm::fail(PM_ERR_INVALID_ARG, "x is NULL");
        if (n == 0) return;
        // Transpose CSR (counting sort over rows in order: each transposed
        // row is sorted by original row index).
        std::vector<int32_t> cp(static_cast<size_t>(n) + 1, 0), ri(nnz);
        std::vector<float> vt(nnz);
        for (int k = 0; k < nnz; ++k) ++cp[static_cast<size_t>(col_idx[k]) + 1];
        for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
        std::vector<int32_t> cur(cp.begin(), cp.end() - 1);
        for (int i = 0; i < m; ++i)
            for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
                const int d = cur[col_idx[k]]++;
                ri[d] = i;
                vt[d] = values[k];
            }
        std::vector<float> xs(x, x + m);
        if (xs.empty()) xs.push_back(0.0f);
        pm::run_spmv(cp.data(), ri.data(), vt.data(), nnz, xs.data(), y, n, m, variant);
    });
}

int pm_launch_pdhg_iteration(float* x, float* y, const float* gradient, const float* ax_bar, const float* b,
                             int32_t n, int32_t m, float tau, float sigma) {
    return pm::guarded([&] {
        using namespace pm;
        if (n < 0 || m < 0) fail(PM_ERR_INVALID_ARG, "negative dimension");
        if (n <= 0 && m <= 0) return;  // pdhg.cu:86
        if (n > 0 && (!x || !gradient)) fail(PM_ERR_INVALID_ARG, "x / gradient is NULL");
        if (m > 0 && (!y || !ax_bar || !b)) fail(PM_ERR_INVALID_ARG, "y / ax_bar / b is NULL");
        PM_HIP(hipSetDevice(0));
        hipStream_t s = nullptr;
        DevBuf<float> dx, dg, dy, da, db;
        if (n > 0) {
            dx.upload(x, n, s);
            dg.upload(gradient, n, s);
            launch_elem(n, FnStep1Primal{dx.get(), dg.get(), tau}, s);
        }
        if (m > 0) {
            dy.upload(y, m, s);
            da.upload(ax_bar, m, s);
            db.upload(b, m, s);
            launch_elem(m, FnStep1Dual{dy.get(), da.get(), db.get(), sigma}, s);
        }
        if (n > 0) dx.download(x, n, s);
        if (m > 0) dy.download(y, m, s);
        PM_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
