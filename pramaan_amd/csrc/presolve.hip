// presolve.hip -- Ruiz equilibration of one LP on the device: the drop-in
// for pramaan::ruiz_scale (reference src/presolve/scaling.cpp:29-150).
//
// Every pass takes the infinity norm of each row and of each column, the
// column norms including |c_j| (scaling.cpp:51-78), turns each norm into
// the factor 1/sqrt(norm) (1 for a zero norm, :80-94), scales A, c, the
// finite bounds and the cumulative factors (:96-137).  All of it is FP64
// with the reference's expressions, so the results equal the reference's
// host loops bit for bit:
//   - a maximum does not depend on the order the entries arrive in, so the
//     column norms come from an atomic maximum over the nonzeros in CSR
//     order, no transpose needed: for non-negative doubles the IEEE bit
//     pattern orders as the value, so the maximum is taken on the 64-bit
//     patterns;
//   - the row norms and the scaling of A run L lanes per row (L from the
//     mean row length), each lane striding over the row's CSR entries so
//     the lanes of a wavefront read contiguous runs; the lanes' maxima are
//     combined by shuffles (exact in any order), and the scaling is
//     v *= rs * cs[j], the reference's product, entry by entry;
//   - c, the bounds and the factors are elementwise.
// The host arrays go to the device once, the passes run there, and the
// scaled arrays come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

namespace pm {
namespace {

constexpr double kBound = 1e30;  // |bound| >= 1e30 is an absent bound (ir.hpp kInfinity)

using Bits = unsigned long long;

dim3 grid_of(int64_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

__global__ __launch_bounds__(kBlock) void k_ones(int n, double* __restrict__ a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = 1.0;
}

// cn[j] <- |c_j| (as bits): the column norm starts from the objective.
__global__ __launch_bounds__(kBlock) void k_col_norm_init(int n, const double* __restrict__ c, Bits* __restrict__ cn) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j < n) cn[j] = static_cast<Bits>(__double_as_longlong(fabs(c[j])));
}

// cn[col] <- max(cn[col], |v_k|) over every nonzero.
__global__ __launch_bounds__(kBlock) void k_col_norm(int64_t nnz, const int* __restrict__ ci,
                                                     const double* __restrict__ v, Bits* __restrict__ cn) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; k < nnz; k += stride) {
        const Bits a = static_cast<Bits>(__double_as_longlong(fabs(v[k])));
        const int j = ci[k];
        // a stale read only costs an atomic that changes nothing
        if (a > cn[j]) atomicMax(cn + j, a);
    }
}

// Column factors, then c, the finite variable bounds and col_scale.
__global__ __launch_bounds__(kBlock) void k_col_apply(int n, const Bits* __restrict__ cn, double* __restrict__ cs,
                                                      double* __restrict__ c, double* __restrict__ lo,
                                                      double* __restrict__ hi, double* __restrict__ acc) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double norm = __longlong_as_double(static_cast<long long>(cn[j]));
    const double f = norm > 0.0 ? 1.0 / sqrt(norm) : 1.0;
    cs[j] = f;
    c[j] *= f;
    if (lo[j] > -kBound) lo[j] = lo[j] / f;
    if (hi[j] < kBound) hi[j] = hi[j] / f;
    acc[j] *= f;
}

// rs[i] <- 1/sqrt(max_k |v_k|) over row i (1 for a zero row), L lanes per
// row.
template <int L>
__global__ __launch_bounds__(kBlock) void k_row_norm_l(int m, const int* __restrict__ rp, const double* __restrict__ v,
                                                       double* __restrict__ rs) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t / L, lane = t % L;
    const bool on = i < m;
    double a = 0.0;
    if (on)
        for (int k = rp[i] + lane; k < rp[i + 1]; k += L) a = fmax(a, fabs(v[k]));
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, L));
    if (on && lane == 0) rs[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// A's rows (v *= rs * cs[col]), L lanes per row; the finite row bounds and
// row_scale by the row's first lane.
template <int L>
__global__ __launch_bounds__(kBlock) void k_row_apply(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                                      double* __restrict__ v, const double* __restrict__ rs,
                                                      const double* __restrict__ cs, double* __restrict__ lo,
                                                      double* __restrict__ hi, double* __restrict__ acc) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t / L, lane = t % L;
    if (i >= m) return;
    const double r = rs[i];
    for (int k = rp[i] + lane; k < rp[i + 1]; k += L) v[k] *= r * cs[ci[k]];
    if (lane != 0) return;
    if (lo[i] > -kBound) lo[i] = lo[i] * r;
    if (hi[i] < kBound) hi[i] = hi[i] * r;
    acc[i] *= r;
}

// Lanes per row for a mean row length: about two entries per lane.
int row_lanes(double mean) {
    int l = 1;
    while (l < 64 && 2.0 * l < mean) l *= 2;
    return l;
}

template <int L>
void row_passes(hipStream_t s, int m, const int* rp, const int* ci, double* v, double* rs, const double* cs,
                double* lo, double* hi, double* acc, bool norm) {
    const dim3 g(static_cast<unsigned>((static_cast<int64_t>(m) * L + kBlock - 1) / kBlock)), b(kBlock);
    if (norm)
        hipLaunchKernelGGL(k_row_norm_l<L>, g, b, 0, s, m, rp, v, rs);
    else
        hipLaunchKernelGGL(k_row_apply<L>, g, b, 0, s, m, rp, ci, v, rs, cs, lo, hi, acc);
}

void rows(int lanes, hipStream_t s, int m, const int* rp, const int* ci, double* v, double* rs, const double* cs,
          double* lo, double* hi, double* acc, bool norm) {
    switch (lanes) {
        case 1: return row_passes<1>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
        case 2: return row_passes<2>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
        case 4: return row_passes<4>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
        case 8: return row_passes<8>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
        case 16: return row_passes<16>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
        case 32: return row_passes<32>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
        default: return row_passes<64>(s, m, rp, ci, v, rs, cs, lo, hi, acc, norm);
    }
}

void check_inputs(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, const double* v, const double* c,
                  const double* rl, const double* ru, const double* vl, const double* vu, int32_t passes,
                  const double* row_scale, const double* col_scale) {
    if (m < 0 || n < 0) fail(PM_ERR_INVALID_ARG, "negative dimension");
    if (passes < 0) fail(PM_ERR_INVALID_ARG, "negative number of passes");
    if (m > 0 && (!rp || !rl || !ru || !row_scale)) fail(PM_ERR_INVALID_ARG, "row_ptr / row bounds / row_scale is NULL");
    if (n > 0 && (!c || !vl || !vu || !col_scale)) fail(PM_ERR_INVALID_ARG, "c / variable bounds / col_scale is NULL");
    if (m == 0) return;
    if (rp[0] != 0) fail(PM_ERR_INVALID_ARG, "row_ptr[0] != 0");
    for (int32_t i = 0; i < m; ++i)
        if (rp[i + 1] < rp[i]) fail(PM_ERR_INVALID_ARG, "row_ptr decreases at row " + std::to_string(i));
    const int32_t nnz = rp[m];
    if (nnz > 0 && (!ci || !v)) fail(PM_ERR_INVALID_ARG, "col_idx / values is NULL");
    for (int32_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= n) fail(PM_ERR_INVALID_ARG, "column index out of range at entry " + std::to_string(k));
}

void ruiz_device(int32_t m, int32_t n, const int32_t* rp, const int32_t* ci, double* v, double* c, double* rl,
                 double* ru, double* vl, double* vu, int32_t passes, double* row_scale, double* col_scale) {
    for (int32_t i = 0; i < m; ++i) row_scale[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) col_scale[j] = 1.0;
    if (m == 0 || n == 0 || passes == 0) return;  // scaling.cpp:37
    PM_HIP(hipSetDevice(0));
    hipStream_t s = nullptr;
    const int64_t nnz = rp[m];
    DevBuf<int> d_rp, d_ci;
    DevBuf<double> d_v, d_c, d_rl, d_ru, d_vl, d_vu, d_dr, d_dc, d_rs, d_cs;
    DevBuf<Bits> d_cn(static_cast<size_t>(n));
    d_rp.upload(rp, static_cast<size_t>(m) + 1, s);
    d_ci.upload(ci, static_cast<size_t>(nnz), s);
    d_v.upload(v, static_cast<size_t>(nnz), s);
    d_c.upload(c, n, s);
    d_rl.upload(rl, m, s);
    d_ru.upload(ru, m, s);
    d_vl.upload(vl, n, s);
    d_vu.upload(vu, n, s);
    d_dr.resize(m);
    d_dc.resize(n);
    d_rs.resize(m);
    d_cs.resize(n);
    const dim3 b(kBlock);
    hipLaunchKernelGGL(k_ones, grid_of(m), b, 0, s, m, d_dr.get());
    hipLaunchKernelGGL(k_ones, grid_of(n), b, 0, s, n, d_dc.get());
    // one nonzero per thread per step of a grid-stride loop
    const int64_t nz_blocks = nnz > 0 ? std::min<int64_t>((nnz + kBlock - 1) / kBlock, 16384) : 0;
    const int lanes = row_lanes(static_cast<double>(nnz) / m);
    for (int32_t pass = 0; pass < passes; ++pass) {
        rows(lanes, s, m, d_rp.get(), nullptr, d_v.get(), d_rs.get(), nullptr, nullptr, nullptr, nullptr, true);
        hipLaunchKernelGGL(k_col_norm_init, grid_of(n), b, 0, s, n, d_c.get(), d_cn.get());
        if (nz_blocks > 0)
            hipLaunchKernelGGL(k_col_norm, dim3(static_cast<unsigned>(nz_blocks)), b, 0, s, nnz, d_ci.get(),
                               d_v.get(), d_cn.get());
        hipLaunchKernelGGL(k_col_apply, grid_of(n), b, 0, s, n, d_cn.get(), d_cs.get(), d_c.get(), d_vl.get(),
                           d_vu.get(), d_dc.get());
        rows(lanes, s, m, d_rp.get(), d_ci.get(), d_v.get(), d_rs.get(), d_cs.get(), d_rl.get(), d_ru.get(),
             d_dr.get(), false);
        PM_LAUNCHED();
    }
    d_v.download(v, static_cast<size_t>(nnz), s);
    d_c.download(c, n, s);
    d_rl.download(rl, m, s);
    d_ru.download(ru, m, s);
    d_vl.download(vl, n, s);
    d_vu.download(vu, n, s);
    d_dr.download(row_scale, m, s);
    d_dc.download(col_scale, n, s);
    PM_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace pm

extern "C" {

int pm_ruiz_scale(int32_t m, int32_t n, const int32_t* row_ptr, const int32_t* col_idx, double* values, double* c,
                  double* row_lower, double* row_upper, double* var_lower, double* var_upper, int32_t passes,
                  double* row_scale, double* col_scale) {
    return pm::guarded([&] {
        pm::check_inputs(m, n, row_ptr, col_idx, values, c, row_lower, row_upper, var_lower, var_upper, passes,
                         row_scale, col_scale);
        pm::ruiz_device(m, n, row_ptr, col_idx, values, c, row_lower, row_upper, var_lower, var_upper, passes,
                        row_scale, col_scale);
    });
}

}  // extern "C"
