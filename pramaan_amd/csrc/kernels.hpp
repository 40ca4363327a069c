// kernels.hpp -- gfx950 device code of the PDHG engine.
//
// Layout in HBM (one LP): A as CSR (row_ptr int32[m+1], col int32[nnz],
// val float[nnz]) and A^T as a second CSR built once at load (col_ptr
// int32[n+1], row int32[nnz], val float[nnz]), so both products are
// row-parallel gathers with no atomics.  The reference computes A^T y by a
// per-row atomicAdd scatter (src/gpu/pdhg_kernels.cuh:42-54); the
// transpose makes the result deterministic and the stream coalesced.
//
// Every SpMV is one templated row kernel: L lanes of a 64-wide wavefront
// share a row (L in 1..64), walk it with stride L (four loads in flight per
// lane), reduce with a butterfly of wavefront shuffles and hand the row sum
// to an epilogue functor.  The PDHG projections, extrapolation and Halpern
// averaging are those epilogues, so one iteration is two launches that each
// stream the matrix once.  All kernels are built with -ffp-contract=off:
// fused multiply-adds are written explicitly (fma), everything else rounds
// exactly as written, which lets the sequential variant reproduce a CPU
// loop bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pm {

constexpr int kBlock = 256;
constexpr float kFltInf = 1e30f;   // FP32 infinity sentinel (pdhg_solver.cu:200)
constexpr double kInf = 1e30;      // ModelIR kInfinity (ir.hpp:47)

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  Remap so each XCD owns one contiguous range of rows: rows
// that share gathered vector entries then share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    if (nb < 16) return b;
    const int x = b & 7, q = b >> 3, per = nb >> 3, rem = nb & 7;
    return x * per + (x < rem ? x : rem) + q;
}

template <int L, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
    return v;
}

template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double fma_t<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Row sum of row `row` by the L lanes of its group (result valid on every
// lane of the group).  EXACT (L == 1 only): sequential, unfused, in CSR
// order -- the reference's csr_spmv_kernel loop (pdhg_kernels.cuh:31-35).
template <int L, bool EXACT, class AT, class VT, class XT>
__device__ __forceinline__ AT row_dot(int beg, int end, int lane, const int* __restrict__ ci,
                                      const VT* __restrict__ v, const XT* __restrict__ x) {
    AT s = AT(0);
    if constexpr (EXACT) {
        for (int k = beg; k < end; ++k) s = s + static_cast<AT>(v[k]) * static_cast<AT>(x[ci[k]]);
    } else {
        int k = beg + lane;
        for (; k + 3 * L < end; k += 4 * L) {
            const int c0 = ci[k], c1 = ci[k + L], c2 = ci[k + 2 * L], c3 = ci[k + 3 * L];
            const AT v0 = static_cast<AT>(v[k]), v1 = static_cast<AT>(v[k + L]);
            const AT v2 = static_cast<AT>(v[k + 2 * L]), v3 = static_cast<AT>(v[k + 3 * L]);
            const AT x0 = static_cast<AT>(x[c0]), x1 = static_cast<AT>(x[c1]);
            const AT x2 = static_cast<AT>(x[c2]), x3 = static_cast<AT>(x[c3]);
            s = fma_t<AT>(v0, x0, s);
            s = fma_t<AT>(v1, x1, s);
            s = fma_t<AT>(v2, x2, s);
            s = fma_t<AT>(v3, x3, s);
        }
        for (; k < end; k += L) s = fma_t<AT>(static_cast<AT>(v[k]), static_cast<AT>(x[ci[k]]), s);
        s = group_sum<L>(s);
    }
    return s;
}

// y-side row kernel: epi(row, sum) on lane 0 of each row group.
template <int L, bool EXACT, class AT, class VT, class XT, class Epi>
__global__ __launch_bounds__(kBlock) void k_rows(int nrows, const int* __restrict__ rp,
                                                 const int* __restrict__ ci,
                                                 const VT* __restrict__ v,
                                                 const XT* __restrict__ x, Epi epi) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int row = (b * kBlock + static_cast<int>(threadIdx.x)) / L;
    const int lane = static_cast<int>(threadIdx.x) & (L - 1);
    if (row >= nrows) return;  // whole row groups leave together
    const AT s = row_dot<L, EXACT, AT>(rp[row], rp[row + 1], lane, ci, v, x);
    if (lane == 0) epi(row, s);
}

// Deterministic block reduction of NS sums followed by NM maxima.  Every
// thread of the block must call it.  Writes one record per block.
template <int NS, int NM>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NS + NM], double* __restrict__ out,
                                                   int block_slot) {
    __shared__ double red[kBlock / 64][NS + NM];
#pragma unroll
    for (int j = 0; j < NS + NM; ++j) {
        double a = acc[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(a, o, 64);
            if (j < NS) a += t;
            else a = (t > a) ? t : a;
        }
        acc[j] = a;
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 0; j < NS + NM; ++j) red[wave][j] = acc[j];
    }
    __syncthreads();
    if (threadIdx.x < NS + NM) {
        const int j = threadIdx.x;
        double a = red[0][j];
        for (int w = 1; w < kBlock / 64; ++w) {
            if (j < NS) a += red[w][j];
            else a = (red[w][j] > a) ? red[w][j] : a;
        }
        out[static_cast<size_t>(block_slot) * (NS + NM) + j] = a;
    }
}

// Row kernel whose epilogue contributes to block partial sums/maxima.
// Epi::operator()(row, sum, acc) adds its terms into acc[NS + NM].
template <int L, bool EXACT, class AT, class VT, class XT, class Epi>
__global__ __launch_bounds__(kBlock) void k_rows_reduce(int nrows, const int* __restrict__ rp,
                                                        const int* __restrict__ ci,
                                                        const VT* __restrict__ v,
                                                        const XT* __restrict__ x, Epi epi,
                                                        double* __restrict__ partials) {
    constexpr int NA = Epi::NS + Epi::NM;
    double acc[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) acc[j] = 0.0;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int row = (b * kBlock + static_cast<int>(threadIdx.x)) / L;
    const int lane = static_cast<int>(threadIdx.x) & (L - 1);
    if (row < nrows) {
        const AT s = row_dot<L, EXACT, AT>(rp[row], rp[row + 1], lane, ci, v, x);
        if (lane == 0) epi(row, s, acc);
    }
    block_reduce_store<Epi::NS, Epi::NM>(acc, partials, b);
}

// Elementwise kernel: f(i, acc) for i < n, then the same block reduction.
template <class F>
__global__ __launch_bounds__(kBlock) void k_elem_reduce(int n, F f, double* __restrict__ partials) {
    constexpr int NA = F::NS + F::NM;
    double acc[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) acc[j] = 0.0;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) f(i, acc);
    block_reduce_store<F::NS, F::NM>(acc, partials, blockIdx.x);
}

// Sums (first ns) and maxima (next nm) of `nblocks` records of width
// ns+nm, in a fixed order: one block, thread j owns blocks j, j+256, ...
// (A template only so that translation units that do not launch it do not
// instantiate it.)
template <int = 0>
__global__ __launch_bounds__(kBlock) void k_final_reduce(const double* __restrict__ partials, int nblocks, int ns,
                                                         int nm, double* __restrict__ out) {
    __shared__ double red[kBlock];
    const int w = ns + nm;
    for (int j = 0; j < w; ++j) {
        const bool is_sum = j < ns;
        double a = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) {
            const double t = partials[static_cast<size_t>(b) * w + j];
            if (is_sum) a += t;
            else a = (t > a) ? t : a;
        }
        red[threadIdx.x] = a;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (static_cast<int>(threadIdx.x) < s) {
                const double t = red[threadIdx.x + s];
                if (is_sum) red[threadIdx.x] += t;
                else red[threadIdx.x] = (t > red[threadIdx.x]) ? t : red[threadIdx.x];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = red[0];
        __syncthreads();
    }
}

template <class F>
__global__ __launch_bounds__(kBlock) void k_elem(int n, F f) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) f(i);
}

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------

template <class T>
struct EpiStore {
    T* y;
    template <class AT>
    __device__ void operator()(int i, AT s) const { y[i] = static_cast<T>(s); }
};

// Sliced ELL row kernel of one LP (rows of A or of A^T): one wavefront per
// slice of 64 rows, lane l owning row srow[s * 64 + l] (-1: a padding
// lane).  The slice's entries are stored column by column, entry j of the
// lane's row at sptr[s] + j * 64 + l, so each load of the matrix is one
// 256-byte coalesced access and the lane walks its row with no row pointer
// and no cross-lane reduction; padding entries hold value 0 and a column
// of their row.  Four entries (eight loads) are in flight per lane.
template <class Epi>
__global__ __launch_bounds__(kBlock) void k_sell(int nslices, const int* __restrict__ sptr,
                                                 const int* __restrict__ slen, const int* __restrict__ srow,
                                                 const int* __restrict__ col, const float* __restrict__ val,
                                                 const float* __restrict__ x, Epi epi) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int s = __builtin_amdgcn_readfirstlane(b * (kBlock / 64) + static_cast<int>(threadIdx.x) / 64);
    if (s >= nslices) return;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int len = __builtin_amdgcn_readfirstlane(slen[s]);
    const int base = __builtin_amdgcn_readfirstlane(sptr[s]);
    const int* __restrict__ c = col + base + lane;
    const float* __restrict__ v = val + base + lane;
    float acc = 0.0f;
    int j = 0;
    for (; j + 3 < len; j += 4, c += 256, v += 256) {
        const int c0 = c[0], c1 = c[64], c2 = c[128], c3 = c[192];
        const float v0 = v[0], v1 = v[64], v2 = v[128], v3 = v[192];
        const float x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc = __builtin_fmaf(v0, x0, acc);
        acc = __builtin_fmaf(v1, x1, acc);
        acc = __builtin_fmaf(v2, x2, acc);
        acc = __builtin_fmaf(v3, x3, acc);
    }
    for (; j < len; ++j, c += 64, v += 64) acc = __builtin_fmaf(v[0], x[c[0]], acc);
    const int r = srow[s * 64 + lane];
    if (r >= 0) epi(r, acc);
}

// Column-blocked product: the launches of one product walk the column
// blocks in order; each adds its row sums to acc[] (the first one starts
// it), and the last hands the row total to the inner epilogue.
template <class Inner>
struct EpiBlock {
    float* acc;
    Inner inner;
    int first, last;
    template <class AT>
    __device__ void operator()(int i, AT s) const {
        float t = static_cast<float>(s);
        if (!first) t = acc[i] + t;
        if (last) inner(i, t);
        else acc[i] = t;
    }
};

// Reference dual update fused into A x-bar (dual_update_kernel,
// pdhg_kernels.cuh:72-89): w = y/sigma + Ax; y = sigma (w - clamp(w,l,u)),
// rows with sigma <= 0 untouched.
struct EpiDualRef {
    const float* sigma;
    const float* lo;
    const float* hi;
    float* y;
    __device__ void operator()(int i, float ax) const {
        const float sig = sigma[i];
        if (sig <= 0.0f) return;
        const float w = y[i] / sig + ax;
        const float wp = fminf(fmaxf(w, lo[i]), hi[i]);
        y[i] = sig * (w - wp);
    }
};

// Reference primal update fused into A^T y (primal_update_kernel,
// pdhg_kernels.cuh:95-109).
struct EpiPrimalRef {
    const float* tau;
    const float* c;
    const float* lo;
    const float* hi;
    float* x;
    float* xbar;
    __device__ void operator()(int j, float aty) const {
        const float x_prev = x[j];
        const float x_new = fminf(fmaxf(x_prev - tau[j] * (c[j] + aty), lo[j]), hi[j]);
        xbar[j] = 2.0f * x_new - x_prev;
        x[j] = x_new;
    }
};

// Reflected Halpern PDHG, primal half (fused into A^T y):
//   xt = clamp(x - tau (c + A^T y), l, u)           (the PDHG step T)
//   xe = 2 xt - x                                    (extrapolation, feeds A xe)
//   x  = lam ((1+rho) xt - rho x) + (1-lam) x0       (reflected Halpern step)
// xt_out (may be null) keeps T(z) for the KKT check.
struct EpiPrimalHal {
    float tau, lam, rho;
    const float* c;
    const float* lo;
    const float* hi;
    const float* x0;
    float* x;
    float* xe;
    float* xt_out;
    // The operands of column j, loadable before its product is known.
    struct Pre {
        float xk, c, lo, hi, x0;
    };
    __device__ Pre load(int j) const { return Pre{x[j], c[j], lo[j], hi[j], x0[j]}; }
    __device__ void apply(int j, float aty, const Pre& p) const {
        const float xt = fminf(fmaxf(p.xk - tau * (p.c + aty), p.lo), p.hi);
        xe[j] = 2.0f * xt - p.xk;
        x[j] = lam * ((1.0f + rho) * xt - rho * p.xk) + (1.0f - lam) * p.x0;
        if (xt_out) xt_out[j] = xt;
    }
    __device__ void operator()(int j, float aty) const { apply(j, aty, load(j)); }
};

// Reflected Halpern PDHG, dual half (fused into A xe).  The row prox is
// that of f(s) = indicator_[l,u](s) + g s: with w = y/sigma + A xe,
//   yt = sigma (w - clamp(w - g/sigma, l, u)),
// which for g = 0 is the reference's dual update; g != 0 carries the dual
// base point of an iterative-refinement round.
struct EpiDualHal {
    float sigma, lam, rho;
    const float* lo;
    const float* hi;
    const float* g;  // may be null
    const float* y0;
    float* y;
    float* yt_out;
    // The operands of row i, loadable before its product is known.
    struct Pre {
        float yk, lo, hi, g, y0;
    };
    __device__ Pre load(int i) const { return Pre{y[i], lo[i], hi[i], g ? g[i] : 0.0f, y0[i]}; }
    __device__ void apply(int i, float ax, const Pre& p) const {
        const float w = p.yk / sigma + ax;
        const float sh = g ? w - p.g / sigma : w;
        const float yt = sigma * (w - fminf(fmaxf(sh, p.lo), p.hi));
        y[i] = lam * ((1.0f + rho) * yt - rho * p.yk) + (1.0f - lam) * p.y0;
        if (yt_out) yt_out[i] = yt;
    }
    __device__ void operator()(int i, float ax) const { apply(i, ax, load(i)); }
};

// The same updates with the step sizes, the dual shift's presence and the
// Halpern weight read from device memory when the kernel runs: dp[0] tau,
// dp[1] sigma, dp[2] 1 with a dual shift (else 0), dp[3 + pos] the weight
// of position pos of a chunk.  One captured chunk of launches then serves
// every chunk of a solve; the arithmetic is that of EpiPrimalHal /
// EpiDualHal with rho = 1.
struct EpiPrimalHalDev {
    const float* dp;
    int pos;
    const float* c;
    const float* lo;
    const float* hi;
    const float* x0;
    float* x;
    float* xe;
    float* xt_out;
    __device__ void operator()(int j, float aty) const {
        EpiPrimalHal{dp[0], dp[3 + pos], 1.0f, c, lo, hi, x0, x, xe, xt_out}(j, aty);
    }
};

struct EpiDualHalDev {
    const float* dp;
    int pos;
    const float* lo;
    const float* hi;
    const float* g;
    const float* y0;
    float* y;
    float* yt_out;
    __device__ void operator()(int i, float ax) const {
        EpiDualHal{dp[1], dp[3 + pos], 1.0f, lo, hi, dp[2] != 0.0f ? g : nullptr, y0, y, yt_out}(i, ax);
    }
};

// ---------------------------------------------------------------------------
// FP64 checks
// ---------------------------------------------------------------------------

// Reference primal violation of the rows (ComputeFp64Residuals,
// pdhg_solver.cu:124-130): max over rows of the bound violation of A x.
struct EpiRefRowViol {
    static constexpr int NS = 0, NM = 1;
    const double* lo;
    const double* hi;
    __device__ void operator()(int i, double act, double* acc) const {
        double pv = 0.0;
        if (act < lo[i] && lo[i] - act > pv) pv = lo[i] - act;
        if (act > hi[i] && act - hi[i] > pv) pv = act - hi[i];
        acc[0] = pv;
    }
};

// Reference bound violation of x and projected reduced-cost violation
// (pdhg_solver.cu:131-157), fused into the FP64 A^T y over the FP32 values.
// acc[0] = max bound violation, acc[1] = max dual violation (before the
// division by max(1, ||c||_inf), which the host applies).
template <class XT>
struct EpiRefColDual {
    static constexpr int NS = 0, NM = 2;
    const XT* x;
    const float* c32;
    const double* lo;
    const double* hi;
    double tol;
    __device__ void operator()(int j, double aty, double* acc) const {
        const double xd = static_cast<double>(x[j]);
        double bv = 0.0;
        if (xd < lo[j] && lo[j] - xd > bv) bv = lo[j] - xd;
        if (xd > hi[j] && xd - hi[j] > bv) bv = xd - hi[j];
        const double rc = static_cast<double>(c32[j]) + aty;
        const bool at_lb = xd <= lo[j] + tol;
        const bool at_ub = xd >= hi[j] - tol;
        double viol = 0.0;
        if (!at_lb && !at_ub) viol = fabs(rc);
        else if (at_lb && !at_ub) viol = (-rc > 0.0) ? -rc : 0.0;
        else if (at_ub && !at_lb) viol = (rc > 0.0) ? rc : 0.0;
        acc[0] = bv;
        acc[1] = (viol > 0.0) ? viol : 0.0;  // NaN never wins a max
    }
};

// Relative-KKT row terms of a total iterate (x_tot gathered, y_tot[i]):
//  sums: [0] |r|^2, [1] |D_r r|^2, [2] bhat^2, [3] row dual-sign residual^2,
//        [4] its scaled form, [5] row part of the dual objective
//  max:  [6] |r|
struct EpiKktRow {
    static constexpr int NS = 6, NM = 1;
    const double* lo;
    const double* hi;
    const double* y;
    const double* dr;
    __device__ void operator()(int i, double act, double* acc) const {
        const double l = lo[i], u = hi[i];
        const bool lf = l > -kInf, uf = u < kInf;
        double p = act;
        if (lf && p < l) p = l;
        if (uf && p > u) p = u;
        const double r = act - p;
        const double d = dr[i];
        double bh = 0.0;
        if (lf && fabs(l) > bh) bh = fabs(l);
        if (uf && fabs(u) > bh) bh = fabs(u);
        const double t = -y[i];
        double tp = t;
        if (!lf && tp > 0.0) tp = 0.0;
        if (!uf && tp < 0.0) tp = 0.0;
        const double e = t - tp;
        acc[0] = r * r;
        acc[1] = (d * r) * (d * r);
        acc[2] = bh * bh;
        acc[3] = e * e;
        acc[4] = (e / d) * (e / d);
        acc[5] = tp > 0.0 ? tp * l : (tp < 0.0 ? tp * u : 0.0);
        acc[6] = fabs(r);
    }
};

// Relative-KKT column terms: rc = c + A^T y_tot, lambda = rc projected on
// the signs x's finite bounds allow.
//  sums: [0] |rc - lambda|^2, [1] |D_c (rc - lambda)|^2, [2] c^2,
//        [3] column part of the dual objective, [4] c^T x
//  max:  [5] |rc - lambda|
struct EpiKktCol {
    static constexpr int NS = 5, NM = 1;
    const double* c;
    const double* lo;
    const double* hi;
    const double* x;
    const double* dc;
    __device__ void operator()(int j, double aty, double* acc) const {
        const double l = lo[j], u = hi[j];
        const bool lf = l > -kInf, uf = u < kInf;
        const double cj = c[j];
        const double rc = cj + aty;
        double lam = rc;
        if (!lf && lam > 0.0) lam = 0.0;
        if (!uf && lam < 0.0) lam = 0.0;
        const double e = rc - lam;
        acc[0] = e * e;
        acc[1] = (dc[j] * e) * (dc[j] * e);
        acc[2] = cj * cj;
        acc[3] = lam > 0.0 ? lam * l : (lam < 0.0 ? lam * u : 0.0);
        acc[4] = cj * x[j];
        acc[5] = fabs(e);
    }
};

__device__ __forceinline__ float to_f32_bound(double v) {
    if (v >= 1e29) return kFltInf;
    if (v <= -1e29) return -kFltInf;
    return static_cast<float>(v);
}

// What a round's correction problem is: the LP itself (refinement rounds),
// or one of the two feasibility problems of a polish (Polish in
// control.hpp): the LP with cost 0, or with every finite bound at 0.
enum RoundMode { kRoundLp = 0, kRoundPrimalFeas = 1, kRoundDualFeas = 2 };

// ---------------------------------------------------------------------------
// Merge-path CSR product (after Merrill and Garland's merge-based SpMV).
// The walk that interleaves the row ends rp[1..nrows] with the nonzero
// indices 0..nnz-1 has nrows + nnz steps (a step either consumes a nonzero
// or ends a row); it is cut into tiles of kMergeTile steps, one workgroup
// each, so every workgroup does the same work however long the rows are.
//  * k_merge_partition (once per matrix): prow[b] = rows ended before step
//    b * kMergeTile, by binary search on the walk; the tile then starts at
//    nonzero b * kMergeTile - prow[b].
//  * k_merge_tile: the workgroup stages its nonzeros' products a_k x_ci[k]
//    and its row ends in LDS with coalesced loads; each thread walks
//    kMergeItems steps of the tile (its start found by binary search in
//    LDS); a segmented scan over the threads (keyed by the row each thread
//    stops inside: within each wavefront by shuffles, then across the
//    wavefronts) hands every thread the part of its first row that the
//    threads before it summed, and the rows that end in the tile go through
//    the epilogue in one coalesced pass.
//  * A row of at most kMergeShort entries crosses at most one tile edge; the
//    tile it starts in finishes it (one wavefront sums the entries past the
//    edge), and the next tile skips it.  A longer row that crosses edges
//    leaves its part of each tile it runs through in carry[b] and its part
//    of the tile it ends in in head[b]; k_merge_fix takes the list of such
//    rows (built once per matrix), one wavefront per row, adds the parts in
//    tile order and applies the epilogue.
// Fixed summation order, no atomics: the results repeat bit for bit.
constexpr int kMergeItems = 8;
constexpr int kMergeTile = kBlock * kMergeItems;
constexpr int kMergeShort = 128;
constexpr int kWave = 64;

template <int TILE>
__global__ __launch_bounds__(kBlock) void k_merge_partition(int nrows, int nnz, int ntiles,
                                                            const int* __restrict__ rp, int* __restrict__ prow) {
    const int b = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    if (b > ntiles) return;
    const int64_t total = static_cast<int64_t>(nrows) + nnz;
    const int64_t d = static_cast<int64_t>(b) * TILE < total ? static_cast<int64_t>(b) * TILE : total;
    // rows ended before step d: the first r with rp[r + 1] > d - r - 1
    int64_t lo = d - nnz > 0 ? d - nnz : 0, hi = d < nrows ? d : nrows;
    while (lo < hi) {
        const int64_t piv = (lo + hi) >> 1;
        if (rp[piv + 1] <= d - piv - 1) lo = piv + 1;
        else hi = piv;
    }
    prow[b] = static_cast<int>(lo);
}

template <class Epi>
__global__ __launch_bounds__(kBlock) void k_merge_tile(int nrows, int nnz, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const float* __restrict__ v,
                                                       const float* __restrict__ x, const int* __restrict__ prow,
                                                       float* __restrict__ carry, float* __restrict__ head, Epi epi) {
    __shared__ float s_prod[kMergeTile];
    __shared__ int s_end[kMergeTile + 1];
    __shared__ float s_sum[kMergeTile];
    __shared__ float s_cv[kBlock];
    __shared__ int s_cr[kBlock];
    __shared__ float s_wv[kBlock / kWave];
    __shared__ int s_wr[kBlock / kWave], s_wf[kBlock / kWave];
    const int b = static_cast<int>(blockIdx.x), tid = static_cast<int>(threadIdx.x);
    const int lane = tid & (kWave - 1), w = tid / kWave;
    const int64_t total = static_cast<int64_t>(nrows) + nnz;
    const int64_t d0 = static_cast<int64_t>(b) * kMergeTile;
    const int64_t d1 = d0 + kMergeTile < total ? d0 + kMergeTile : total;
    const int r0 = prow[b], r1 = prow[b + 1];
    const int k0 = static_cast<int>(d0 - r0), k1 = static_cast<int>(d1 - r1);
    const int nr = r1 - r0, nk = k1 - k0;  // rows ended in the tile, nonzeros consumed
    for (int j = tid; j < nk; j += kBlock) s_prod[j] = v[k0 + j] * x[ci[k0 + j]];
    for (int j = tid; j < nr; j += kBlock) s_end[j] = rp[r0 + 1 + j] - k0;
    if (tid == 0) s_end[nr] = 0x7fffffff;  // the row the tile stops inside does not end here
    __syncthreads();
    const int len = nr + nk;
    const int dl = tid * kMergeItems < len ? tid * kMergeItems : len;
    const int de = dl + kMergeItems < len ? dl + kMergeItems : len;
    int lo = dl - nk > 0 ? dl - nk : 0, hi = dl < nr ? dl : nr;
    while (lo < hi) {
        const int piv = (lo + hi) >> 1;
        if (s_end[piv] <= dl - piv - 1) lo = piv + 1;
        else hi = piv;
    }
    const int first = lo;
    int row = lo, k = dl - lo;
    float s = 0.0f;
    for (int d = dl; d < de; ++d) {
        if (k < s_end[row]) {
            s += s_prod[k];
            ++k;
        } else {
            s_sum[row] = s;
            s = 0.0f;
            ++row;
        }
    }
    // inclusive scan of the carries, segmented by the row they belong to
    // (threads stopping inside one row are consecutive): within the
    // wavefront, then the runs that reach back into earlier wavefronts
    float cv = s;
    for (int off = 1; off < kWave; off <<= 1) {
        const float vu = __shfl_up(cv, off);
        const int ru = __shfl_up(row, off);
        if (lane >= off && ru == row) cv += vu;
    }
    const int row_l0 = __shfl(row, 0);
    if (lane == kWave - 1) {
        s_wv[w] = cv;
        s_wr[w] = row;
        s_wf[w] = row_l0;
    }
    __syncthreads();
    if (row == row_l0) {  // this thread's run starts at the wavefront's first lane
        for (int w2 = w - 1; w2 >= 0 && s_wr[w2] == row; --w2) {
            cv += s_wv[w2];
            if (s_wf[w2] != row) break;
        }
    }
    s_cv[tid] = cv;
    s_cr[tid] = row;
    __syncthreads();
    if (row > first && tid > 0 && s_cr[tid - 1] == first) s_sum[first] += s_cv[tid - 1];
    __syncthreads();
    // the first row, when it came in from the previous tile: a short one was
    // finished there, a long one leaves its part here for k_merge_fix
    const bool head_in = nr > 0 && k0 > rp[r0];
    const bool head_long = head_in && rp[r0 + 1] - rp[r0] > kMergeShort;
    for (int j = tid; j < nr; j += kBlock) {
        if (j == 0 && head_in) {
            if (head_long) head[b] = s_sum[0];
        } else {
            epi(r0 + j, s_sum[j]);
        }
    }
    // the row the tile stops inside
    if (r1 < nrows && k1 > rp[r1]) {
        const int kend = rp[r1 + 1];
        if (kend - rp[r1] > kMergeShort) {
            if (tid == kBlock - 1) carry[b] = s_cv[kBlock - 1];
        } else if (w == 0) {
            float e = 0.0f;
            for (int q = k1 + lane; q < kend; q += kWave) e += v[q] * x[ci[q]];
            for (int off = kWave / 2; off > 0; off >>= 1) e += __shfl_xor(e, off);
            if (lane == 0) epi(r1, s_cv[kBlock - 1] + e);
        }
    }
}

// Rows of more than kMergeShort entries that cross tile edges: {row, first
// tile ending inside it, tile it ends in}.  One wavefront per row.
template <class Epi>
__global__ __launch_bounds__(kBlock) void k_merge_fix(int nlong, const int4* __restrict__ longs,
                                                      const float* __restrict__ carry,
                                                      const float* __restrict__ head, Epi epi) {
    const int q = static_cast<int>((blockIdx.x * kBlock + threadIdx.x) / kWave);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    if (q >= nlong) return;
    const int4 L = longs[q];
    float s = 0.0f;
    for (int bb = L.y + lane; bb < L.z; bb += kWave) s += carry[bb];
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) epi(L.x, s + head[L.z]);
}

// Correction-LP column data of a refinement round from FP64 A^T y_base:
// c_hat = D_c DD (c + A^T y_base), bounds DP (l - x_base) / D_c (cost 0 in
// kRoundPrimalFeas, finite bounds 0 in kRoundDualFeas, whose x_base is 0).
struct EpiRoundCol {
    const double* c;
    const double* lo;
    const double* hi;
    const double* xb;
    const double* dc;
    double dp, dd;
    float* c32;
    float* lo32;
    float* hi32;
    int mode = kRoundLp;
    __device__ void operator()(int j, double aty) const {
        c32[j] = mode == kRoundPrimalFeas ? 0.0f : to_f32_bound(dc[j] * dd * (c[j] + aty));
        const bool homog = mode == kRoundDualFeas;
        lo32[j] = lo[j] > -kInf ? (homog ? 0.0f : to_f32_bound(dp * (lo[j] - xb[j]) / dc[j])) : -kFltInf;
        hi32[j] = hi[j] < kInf ? (homog ? 0.0f : to_f32_bound(dp * (hi[j] - xb[j]) / dc[j])) : kFltInf;
    }
};
// Row data: bounds D_r DP (l - A x_base), shift g = -DD y_base / D_r
// (finite bounds 0 in kRoundDualFeas).
struct EpiRoundRow {
    const double* lo;
    const double* hi;
    const double* yb;
    const double* dr;
    double dp, dd;
    float* lo32;
    float* hi32;
    float* g32;
    int mode = kRoundLp;
    __device__ void operator()(int i, double ax) const {
        const bool homog = mode == kRoundDualFeas;
        lo32[i] = lo[i] > -kInf ? (homog ? 0.0f : to_f32_bound(dr[i] * dp * (lo[i] - ax))) : -kFltInf;
        hi32[i] = hi[i] < kInf ? (homog ? 0.0f : to_f32_bound(dr[i] * dp * (hi[i] - ax))) : kFltInf;
        g32[i] = to_f32_bound(-dd * yb[i] / dr[i]);
    }
};

}  // namespace pm
