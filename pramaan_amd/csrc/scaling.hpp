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

// v[k] *= rs[row] * cs[col] over one CSR order; acc[row] *= rs[row].
__global__ __launch_bounds__(kBlock) void k_row_scale(int nrows, const int* __restrict__ rp,
                                                      const int* __restrict__ col, double* __restrict__ v,
                                                      const double* __restrict__ rs, const double* __restrict__ cs,
                                                      double* __restrict__ acc) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrows) return;
    const double r = rs[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) v[k] *= r * cs[col[k]];
    if (acc) acc[i] *= r;
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

// Scales the m x n matrix whose values are in vs (CSR rp / ci) and vst (CSC
// cp / ri), both overwritten with the scaled values; dr[m], dc[n] receive
// the factors; rn[m], cn[n] are workspace.
inline void ruiz_pc_device(hipStream_t s, int passes, int m, int n, const int* rp, const int* ci, const int* cp,
                           const int* ri, double* vs, double* vst, double* dr, double* dc, double* rn, double* cn) {
    const dim3 b(kBlock);
    if (m > 0) hipLaunchKernelGGL(k_fill_f64, grid_for(m), b, 0, s, static_cast<int64_t>(m), dr, 1.0);
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, grid_for(n), b, 0, s, static_cast<int64_t>(n), dc, 1.0);
    for (int pass = 0; pass <= passes; ++pass) {
        const bool sum = pass == passes;  // the last pass is Pock-Chambolle
        if (m > 0) hipLaunchKernelGGL(k_row_norm, grid_for(m), b, 0, s, m, rp, vs, sum, rn);
        if (n > 0) hipLaunchKernelGGL(k_row_norm, grid_for(n), b, 0, s, n, cp, vst, sum, cn);
        if (m > 0 && n > 0) {
            hipLaunchKernelGGL(k_row_scale, grid_for(m), b, 0, s, m, rp, ci, vs, rn, cn, dr);
            hipLaunchKernelGGL(k_row_scale, grid_for(n), b, 0, s, n, cp, ri, vst, cn, rn, dc);
        }
    }
}

}  // namespace
}  // namespace pm
