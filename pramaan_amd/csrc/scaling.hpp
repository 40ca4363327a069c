// scaling.hpp -- Ruiz (infinity-norm passes) + Pock-Chambolle (alpha = 1)
// scaling of A on the device, shared by the single-LP engine and the batch
// engine (reference src/presolve/scaling.cpp:29-150 scales rows and columns
// by 1/sqrt of their infinity norms in the same way; the last pass here is
// Pock-Chambolle, 1/sqrt of the 1-norms).
//
// A is kept in both orders (CSR values v, CSC values vt) and both are
// scaled every pass, so row and column norms are row reductions in one
// order or the other; every norm is a sequential loop in CSR (rows) or CSC
// (columns, rows ascending) order, and every factor and update is the same
// FP64 expression as a plain loop over the entries, so the results are
// reproducible bit for bit on the host.  Every norm and update involves one
// row or column only, so a block-diagonal concatenation of LPs (the batch
// engine) is scaled exactly as each LP alone.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

namespace pm {
namespace {

// out[i] = 1/sqrt(max_k |v_k|) (or of the sum when sum), 1 for an empty or
// zero row.
__global__ __launch_bounds__(kBlock) void k_row_norm(int nrows, const int* __restrict__ rp,
                                                     const double* __restrict__ v, bool sum,
                                                     double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrows) return;
    double a = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) a = sum ? a + fabs(v[k]) : fmax(a, fabs(v[k]));
    out[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// The same infinity norms with the rows of more than kLongRow entries
// walked by the whole wavefront (a maximum is exact in any order, so the
// factors are those of the sequential loop): one lane per row otherwise.
constexpr int kLongRow = 256;

__global__ __launch_bounds__(kBlock) void k_row_norm_max(int nrows, const int* __restrict__ rp,
                                                         const double* __restrict__ v, double* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    int b = 0, e = 0;
    if (i < nrows) {
        b = rp[i];
        e = rp[i + 1];
    }
    const bool lng = e - b > kLongRow;
    double a = 0.0;
    if (!lng)
        for (int k = b; k < e; ++k) a = fmax(a, fabs(v[k]));
    uint64_t mask = __ballot(lng);  // the wavefront's long rows, one at a time
    while (mask) {
        const int src = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        const int bb = __shfl(b, src), ee = __shfl(e, src);
        double mx = 0.0;
        for (int k = bb + lane; k < ee; k += 64) mx = fmax(mx, fabs(v[k]));
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if (lane == src) a = mx;
    }
    if (i < nrows) out[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// v[k] *= rs[row] * cs[col] over one CSR order, L lanes per row striding
// over its entries (each entry is scaled on its own, so the lanes change
// nothing but the access pattern); acc[row] *= rs[row].
template <int L>
__global__ __launch_bounds__(kBlock) void k_row_scale(int nrows, const int* __restrict__ rp,
                                                      const int* __restrict__ col, double* __restrict__ v,
                                                      const double* __restrict__ rs, const double* __restrict__ cs,
                                                      double* __restrict__ acc) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int i = static_cast<int>(t / L), lane = static_cast<int>(t % L);
    if (i >= nrows) return;
    const double r = rs[i];
    for (int k = rp[i] + lane; k < rp[i + 1]; k += L) v[k] *= r * cs[col[k]];
    if (acc && lane == 0) acc[i] *= r;
}

// Lanes per row of k_row_scale for a mean row length: about two entries per
// lane.
inline int scale_lanes(double mean) {
    int l = 1;
    while (l < 32 && 2.0 * l < mean) l *= 2;
    return l;
}

inline void row_scale(hipStream_t s, int lanes, int nrows, const int* rp, const int* col, double* v, const double* rs,
                      const double* cs, double* acc) {
    auto go = [&](auto lc) {
        constexpr int L = decltype(lc)::value;
        const dim3 g(static_cast<unsigned>((static_cast<int64_t>(nrows) * L + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_row_scale<L>, g, dim3(kBlock), 0, s, nrows, rp, col, v, rs, cs, acc);
    };
    switch (lanes) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 16: go(std::integral_constant<int, 16>{}); break;
        default: go(std::integral_constant<int, 32>{}); break;
    }
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) b[i] = static_cast<float>(a[i]);
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(int64_t n, double* __restrict__ a, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

inline dim3 grid_for(int64_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

// Scales the m x n matrix with nnz nonzeros whose values are in vs (CSR
// rp / ci) and vst (CSC cp / ri), both overwritten with the scaled values;
// dr[m], dc[n] receive the factors; rn[m], cn[n] are workspace.
inline void ruiz_pc_device(hipStream_t s, int passes, int m, int n, int64_t nnz, const int* rp, const int* ci,
                           const int* cp, const int* ri, double* vs, double* vst, double* dr, double* dc, double* rn,
                           double* cn) {
    const dim3 b(kBlock);
    const int la = scale_lanes(m > 0 ? static_cast<double>(nnz) / m : 0.0);
    const int lt = scale_lanes(n > 0 ? static_cast<double>(nnz) / n : 0.0);
    if (m > 0) hipLaunchKernelGGL(k_fill_f64, grid_for(m), b, 0, s, static_cast<int64_t>(m), dr, 1.0);
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, grid_for(n), b, 0, s, static_cast<int64_t>(n), dc, 1.0);
    for (int pass = 0; pass <= passes; ++pass) {
        const bool sum = pass == passes;  // the last pass is Pock-Chambolle (sums, in CSR / CSC order)
        if (sum) {
            if (m > 0) hipLaunchKernelGGL(k_row_norm, grid_for(m), b, 0, s, m, rp, vs, true, rn);
            if (n > 0) hipLaunchKernelGGL(k_row_norm, grid_for(n), b, 0, s, n, cp, vst, true, cn);
        } else {
            if (m > 0) hipLaunchKernelGGL(k_row_norm_max, grid_for(m), b, 0, s, m, rp, vs, rn);
            if (n > 0) hipLaunchKernelGGL(k_row_norm_max, grid_for(n), b, 0, s, n, cp, vst, cn);
        }
        if (m > 0 && n > 0) {
            row_scale(s, la, m, rp, ci, vs, rn, cn, dr);
            row_scale(s, lt, n, cp, ri, vst, cn, rn, dc);
        }
    }
}

}  // namespace
}  // namespace pm
