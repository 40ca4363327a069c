// engine.hip -- the device-resident PDHG engine behind the C-ABI.
//
// pm_pdhg_solve replaces gpu::PdhgSolver::solve (reference
// src/gpu/pdhg_solver.cu:169-411).  Two algorithms share the kernels:
//
//  * PM_ALG_REFERENCE restates the reference solver: canonicalisation to
//    FP32 (196-229), diagonal steps 0.95/sum|A| (231-250), x0 = clamp(0)
//    (288-297), dual-then-primal FP32 iterations (319-338), FP64 residual
//    check every check_frequency iterations (340-360), stagnation restarts
//    (362-381), final residuals and FP64 objective (385-410).  The check
//    runs on the device in FP64 and only two numbers cross PCIe.
//
//  * PM_ALG_PDLP: reflected Halpern PDHG on the Ruiz (10 inf-norm passes) +
//    Pock-Chambolle (alpha = 1) scaled LP with constant step 0.99/||A||_2
//    (power iteration), KKT-driven adaptive restarts (0.2 / 0.8 / 0.36), a
//    primal-weight update at each restart, and an FP64 iterative-refinement
//    outer loop: each round solves, in FP32, the correction LP of the FP64
//    residuals of the current point, amplified by 1/KKT.  Every decision
//    uses the FP64 relative KKT error of the original LP; it is also the
//    termination test under PM_TERM_RELATIVE_KKT, while PM_TERM_REFERENCE
//    further requires the reference's pv, dv <= tolerance.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <tuple>

#include "control.hpp"
#include "dispatch.hpp"
#include "scaling.hpp"

namespace pm {

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

// ---------------------------------------------------------------------------
// Elementwise functors
// ---------------------------------------------------------------------------
template <class T>
struct FnFill {
    T* p;
    T v;
    __device__ void operator()(int i) const { p[i] = v; }
};
template <class T>
struct FnCopy {
    T* dst;
    const T* src;
    __device__ void operator()(int i) const { dst[i] = src[i]; }
};
// x = clamp(0, lo, hi) into up to three buffers (pdhg_solver.cu:288-297).
struct FnInitX {
    const float* lo;
    const float* hi;
    float* a;
    float* b;
    float* c;
    __device__ void operator()(int i) const {
        const float v = fmaxf(lo[i], fminf(0.0f, hi[i]));
        a[i] = v;
        if (b) b[i] = v;
        if (c) c[i] = v;
    }
};
// x_tot = clamp(x_base + D_c x_hat / DP, l, u) in FP64.
struct FnTotalX {
    const double* base;
    const double* dc;
    const float* xh;
    double inv;
    const double* lo;
    const double* hi;
    double* out;
    __device__ void operator()(int j) const {
        double v = base[j] + dc[j] * static_cast<double>(xh[j]) * inv;
        if (v < lo[j]) v = lo[j];
        if (v > hi[j]) v = hi[j];
        out[j] = v;
    }
};
// y_tot = y_base + D_r y_hat / DD, projected on the sign its row allows
// (reference sign: no upper bound -> y <= 0, no lower bound -> y >= 0).
struct FnTotalY {
    const double* base;
    const double* dr;
    const float* yh;
    double inv;
    const double* lo;
    const double* hi;
    double* out;
    __device__ void operator()(int i) const {
        double v = base[i] + dr[i] * static_cast<double>(yh[i]) * inv;
        if (!(hi[i] < kInf) && v > 0.0) v = 0.0;
        if (!(lo[i] > -kInf) && v < 0.0) v = 0.0;
        out[i] = v;
    }
};
struct FnNormDiff {  // sums: [0] |a - b|^2
    static constexpr int NS = 1, NM = 0;
    const float* a;
    const float* b;
    __device__ void operator()(int i, double* acc) const {
        const double d = static_cast<double>(a[i]) - static_cast<double>(b[i]);
        acc[0] = d * d;
    }
};
struct FnNormSq {
    static constexpr int NS = 1, NM = 0;
    const float* a;
    __device__ void operator()(int i, double* acc) const {
        const double d = static_cast<double>(a[i]);
        acc[0] = d * d;
    }
};
// v = w / sqrt(ss) with ss read on the device (power iteration).
struct FnNormalize {
    const float* w;
    const double* ss;
    float* v;
    __device__ void operator()(int i) const {
        const double s = ss[0];
        v[i] = s > 0.0 ? static_cast<float>(static_cast<double>(w[i]) / sqrt(s)) : 0.0f;
    }
};

// PDLP's normalized duality gap of a candidate z = (x, y) of the current
// round's FP32 problem, on the ball of radius r = ||z - z0||_omega around
// it (||(dx, dy)||^2_omega = omega |dx|^2 + |dy|^2 / omega): the largest
// Lagrangian gap L(x, y^) - L(x^, y) over the ball, divided by r.  For a
// step scale t the maximiser of the gap minus ||z^ - z||^2_omega / (2 t) is
// one prox step from z with steps t / omega (primal) and t omega (dual);
// its distance grows with t, so the ball's maximiser is that step at the t
// whose distance is r.  One pass evaluates kNdgT values of t at once:
// sums [2k] = the squared distance contribution, [2k + 1] = the gap
// contribution at t[k].  Engine sign: reduced cost c + A^T y, and a row's
// y - g is >= 0 only with a finite upper bound, <= 0 only with a finite
// lower one, g the round's dual shift.
constexpr int kNdgT = 16;
struct NdgTs {
    double t[kNdgT];
};
struct FnNdgX {
    static constexpr int NS = 2 * kNdgT, NM = 0;
    const float* x;    // candidate
    const float* aty;  // A^T y of the candidate's y
    const float* c;
    const float* lo;
    const float* hi;
    double omega;
    NdgTs ts;
    __device__ void operator()(int j, double* acc) const {
        const double xc = x[j], g = static_cast<double>(c[j]) + static_cast<double>(aty[j]);
        const double l = lo[j], u = hi[j];
#pragma unroll
        for (int k = 0; k < kNdgT; ++k) {
            double xh = xc - ts.t[k] / omega * g;
            xh = xh < l ? l : (xh > u ? u : xh);
            const double d = xh - xc;
            acc[2 * k] = omega * d * d;
            acc[2 * k + 1] = -g * d;
        }
    }
};
struct FnNdgY {
    static constexpr int NS = 2 * kNdgT, NM = 0;
    const float* y;   // candidate
    const float* ax;  // A x of the candidate's x
    const float* lo;
    const float* hi;
    const float* gs;  // dual shift, or null
    double omega;
    NdgTs ts;
    // support term of the row: (y - g) u when y - g > 0, (y - g) l when < 0
    __device__ static double supp(double v, double l, double u) { return v > 0.0 ? v * u : (v < 0.0 ? v * l : 0.0); }
    __device__ void operator()(int i, double* acc) const {
        const double yc = y[i], a = ax[i], l = lo[i], u = hi[i], g = gs ? static_cast<double>(gs[i]) : 0.0;
        const double h0 = supp(yc - g, l, u);
#pragma unroll
        for (int k = 0; k < kNdgT; ++k) {
            const double sg = ts.t[k] * omega;
            const double sv = yc / sg + a - g / sg;
            const double yh = g + sg * (sv - (sv < l ? l : (sv > u ? u : sv)));
            const double d = yh - yc;
            acc[2 * k] = d * d / omega;
            acc[2 * k + 1] = a * d - (supp(yh - g, l, u) - h0);
        }
    }
};
// dst = a - b (FP32).
struct FnSubF {
    float* dst;
    const float* a;
    const float* b;
    __device__ void operator()(int i) const { dst[i] = a[i] - b[i]; }
};
// sums: [0] (y - ys)^T (A dx), the interaction term of PDLP's step bound.
struct FnDyAdx {
    static constexpr int NS = 1, NM = 0;
    const float* y;
    const float* ys;
    const float* adx;
    __device__ void operator()(int i, double* acc) const {
        acc[0] = (static_cast<double>(y[i]) - static_cast<double>(ys[i])) * static_cast<double>(adx[i]);
    }
};
struct FnNdgScale {  // sums: [0] |c + A^T y|^2, [1] |A x|^2 over the two sides
    static constexpr int NS = 1, NM = 0;
    const float* a;
    const float* b;  // added to a, or null
    __device__ void operator()(int i, double* acc) const {
        const double v = static_cast<double>(a[i]) + (b ? static_cast<double>(b[i]) : 0.0);
        acc[0] = v * v;
    }
};

// ---------------------------------------------------------------------------
// SpMV autotuning cache: (rows, cols, nnz) -> variant
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_tune_cache;

int heuristic_variant(int nrows, int nnz) {
    const double avg = nrows > 0 ? static_cast<double>(nnz) / nrows : 0.0;
    const int want = static_cast<int>(std::ceil(avg / 3.0));
    int l = 1;
    while (l < want && l < 32) l <<= 1;
    switch (l) {
        case 1: return PM_SPMV_SCALAR;
        case 2: return PM_SPMV_VECTOR2;
        case 4: return PM_SPMV_VECTOR4;
        case 8: return PM_SPMV_VECTOR8;
        case 16: return PM_SPMV_VECTOR16;
        default: return PM_SPMV_VECTOR32;
    }
}

// Sliced ELL copy of one matrix (k_sell): slices of 64 rows, rows sorted
// by length (longest first) inside windows of kSellWindow rows.
constexpr int kSellWindow = 1024;
struct SellMat {
    int slices = 0;
    DevBuf<int> ptr, len, row, col;
    DevBuf<float> val;
    void reset() {
        slices = 0;
        for (auto* b : {&ptr, &len, &row, &col}) b->reset();
        val.reset();
    }
};

}  // namespace pm

using namespace pm;

// ---------------------------------------------------------------------------
// The context
// ---------------------------------------------------------------------------
struct pm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Captured chunk of PDLP iterations (hal_chunk): the launches of
    // chunk_len iterations whose epilogues read their step parameters from
    // chunk_dp (written from chunk_hp before each replay).
    hipGraph_t chunk_graph = nullptr;
    hipGraphExec_t chunk_exec = nullptr;
    int chunk_len = 0;
    // The same for the reference algorithm's iterations (ref_chunk), whose
    // launches take no per-iteration parameters.
    hipGraph_t ref_graph = nullptr;
    hipGraphExec_t ref_exec = nullptr;
    int ref_len = 0;
    bool use_graph = true;
    pm_pdhg_options opts{};
    bool loaded = false;

    int m = 0, n = 0, nnz = 0;
    int sense = 0;
    double offset = 0.0;
    bool nan_cost = false;
    bool exact_f32 = false;
    double c_inf32 = 1.0;  // max(1, ||c32||_inf) for the reference dv

    // Host FP64 model in minimisation form (c negated for a maximisation).
    std::vector<double> c_model, cmin, lr, ur, lx, ux;
    std::vector<double> x0_h, y0_h;
    int warm_accepted = -1;  // last solve: 1 warm start used, 0 replaced by the cold start, -1 none

    // Device structure: A (rp, ci) and A^T (cp, ri).
    DevBuf<int> rp, ci, cp, ri;
    DevBuf<float> va, vat;          // FP32 working values (reference: A, PDLP: scaled A)
    DevBuf<float> oa32, oat32;      // original values, when exactly FP32 (PDLP only)
    DevBuf<double> oa64, oat64;     // original values otherwise
    DevBuf<double> d_cmin, d_lr, d_ur, d_lx, d_ux, d_dr, d_dc;

    // FP32 working vectors.
    DevBuf<float> c32, lr32, ur32, lx32, ux32, sig32, tau32, g32;
    DevBuf<float> chunk_dp;
    PinnedBuf<float> chunk_hp;
    DevBuf<float> x, y, xe, x0, y0, xt, yt, xl, yl, tmp_n, tmp_m;
    // FP64 vectors.
    DevBuf<double> xtot, ytot, xbase, ybase, act64;
    DevBuf<double> xbest, ybest;  // PDLP: the checked candidate of lowest relative KKT error
    // PDLP feasibility polishing: the candidate it starts from, the polished
    // primal, and the main solve's state held meanwhile (FP32 blocks
    // x, x0, xl, c32, lx32, ux32 of n, then y, y0, yl, lr32, ur32, g32 of m;
    // FP64 x_base, y_base).
    DevBuf<double> xsave, ysave, xpol, hold64;
    DevBuf<float> hold32;
    // PDLP restarts on the normalized duality gap (PM_RESTART_GAP): block
    // partials of the multi-t passes.
    bool restart_on_gap = false;
    // PM_STEP_ADAPTIVE: the current step size and the saved iterate of a try
    bool adaptive_step = false;
    double eta_ad = 1.0;
    DevBuf<float> xs, ys, adx;
    DevBuf<double> ndg_part, ndg_red;
    PinnedBuf<double> ndg_h;
    DevBuf<double> partials, red;
    PinnedBuf<double> hred;

    int var_a = PM_SPMV_VECTOR8, var_at = PM_SPMV_VECTOR8;
    double eta = 1.0, omega0 = 1.0, sigma_max = 1.0;

    // Column-blocked copy of the scaled A for the iteration's A x-bar
    // (PDLP): block b holds the entries of every row whose column lies in
    // [b * cols / nblk, (b + 1) * cols / nblk), as a CSR whose row pointers
    // rpb[b * (m + 1) + i] index the block-major cib / vb.  Each block's
    // launch gathers from one slab of x-bar, which then stays in L2.
    int nblk = 0;  // 0: the plain row kernel (var_a) is used
    int blk_var = PM_SPMV_VECTOR4;
    DevBuf<int> rpb, cib;
    DevBuf<float> vb, accb;

    // Sliced ELL copies of the scaled A and A^T (PM_SPMV_SLICED, PDLP):
    // used for the iteration's products when slices > 0.
    SellMat sell_a, sell_at;

    // Merge-path products of the iteration (PM_SPMV_MERGE, PDLP): tile
    // partitions of A and A^T, per-tile carries and head parts.
    bool merge_a = false, merge_at = false;
    DevBuf<int> mpart_a, mpart_at;
    DevBuf<int4> mlong_a, mlong_at;  // long rows crossing tile edges
    int nlong_a = 0, nlong_at = 0;
    DevBuf<float> mcarry, mhead;

    const void* orig_a() const { return exact_f32 ? static_cast<const void*>(oa32.get()) : oa64.get(); }
};

namespace pm {
namespace {

void check_lp(const pm_lp* lp) {
    if (!lp) fail(PM_ERR_INVALID_ARG, "lp is NULL");
    if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) fail(PM_ERR_INVALID_ARG, "negative LP dimension");
    if (lp->n > 0 && (!lp->c || !lp->var_lower || !lp->var_upper))
        fail(PM_ERR_INVALID_ARG, "c / var bounds are NULL");
    if (lp->m > 0 && (!lp->row_lower || !lp->row_upper)) fail(PM_ERR_INVALID_ARG, "row bounds are NULL");
    if (!lp->row_ptr) fail(PM_ERR_INVALID_ARG, "row_ptr is NULL");
    if (lp->nnz > 0 && (!lp->col_idx || !lp->values)) fail(PM_ERR_INVALID_ARG, "col_idx / values are NULL");
    if (lp->row_ptr[0] != 0 || lp->row_ptr[lp->m] != lp->nnz)
        fail(PM_ERR_INVALID_ARG, "row_ptr[0] must be 0 and row_ptr[m] must equal nnz");
    for (int i = 0; i < lp->m; ++i)
        if (lp->row_ptr[i] > lp->row_ptr[i + 1]) fail(PM_ERR_INVALID_ARG, "row_ptr must be non-decreasing");
    for (int k = 0; k < lp->nnz; ++k)
        if (lp->col_idx[k] < 0 || lp->col_idx[k] >= lp->n) fail(PM_ERR_INVALID_ARG, "col_idx entry out of bounds");
    for (int j = 0; j < lp->n; ++j)
        if (lp->var_lower[j] > lp->var_upper[j]) fail(PM_ERR_INVALID_ARG, "var_lower > var_upper");
    for (int i = 0; i < lp->m; ++i)
        if (lp->row_lower[i] > lp->row_upper[i]) fail(PM_ERR_INVALID_ARG, "row_lower > row_upper");
}

void sync(pm_ctx* c) { PM_HIP(hipStreamSynchronize(c->stream)); }

// Transposes (rp, ci, v) of an m x n CSR into (cp, ri, vt): counting sort
// over rows in increasing order, so each transposed row keeps its entries
// sorted (the reference's CSRMatrix::transpose, sparse_matrix.cpp:13-57).
template <class T>
void host_transpose(int m, int n, const int* rp, const int* ci, const T* v, std::vector<int>& cp,
                    std::vector<int>& ri, std::vector<T>& vt, std::vector<int64_t>* perm) {
    const int nnz = rp[m];
    cp.assign(static_cast<size_t>(n) + 1, 0);
    for (int k = 0; k < nnz; ++k) ++cp[static_cast<size_t>(ci[k]) + 1];
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    std::vector<int> cur(cp.begin(), cp.end() - 1);
    ri.resize(nnz);
    if (v) vt.resize(nnz);
    if (perm) perm->resize(nnz);
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int d = cur[ci[k]]++;
            ri[d] = i;
            if (v) vt[d] = v[k];
            if (perm) (*perm)[d] = k;
        }
}

// dst[q] = src[perm[q]]: the block-major copy of FP32 values.
__global__ __launch_bounds__(kBlock) void k_gather_f32(int64_t n, const int* __restrict__ perm,
                                                       const float* __restrict__ src, float* __restrict__ dst) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (q < n) dst[q] = src[perm[q]];
}

// A x-bar through the column blocks with epilogue epi on the row totals.
template <class Epi>
void launch_blocked(pm_ctx* c, int nblk, int var, const float* x, const Epi& epi, hipStream_t s) {
    const int m = c->m;
    for (int b = 0; b < nblk; ++b)
        launch_rows<float>(var, m, c->rpb.get() + static_cast<size_t>(b) * (m + 1), c->cib.get(), c->vb.get(), x,
                           EpiBlock<Epi>{c->accb.get(), epi, b == 0, b == nblk - 1}, s);
}

template <class Epi>
void launch_sell_mat(const SellMat& S, const float* x, const Epi& epi, hipStream_t s) {
    launch_sell(S.slices, S.ptr.get(), S.len.get(), S.row.get(), S.col.get(), S.val.get(), x, epi, s);
}

// The dual half of a PDLP iteration: A x-bar with the fused dual update.
template <class Epi>
void launch_dual_half(pm_ctx* c, const Epi& ed, hipStream_t s) {
    if (c->nblk > 0) launch_blocked(c, c->nblk, c->blk_var, c->xe.get(), ed, s);
    else if (c->merge_a)
        launch_merge(c->m, c->nnz, c->rp.get(), c->ci.get(), c->va.get(), c->xe.get(), ed, c->mpart_a.get(),
                     c->mlong_a.get(), c->nlong_a, c->mcarry.get(), c->mhead.get(), s);
    else if (c->sell_a.slices > 0) launch_sell_mat(c->sell_a, c->xe.get(), ed, s);
    else launch_rows<float>(c->var_a, c->m, c->rp.get(), c->ci.get(), c->va.get(), c->xe.get(), ed, s);
}

// The primal half of a PDLP iteration: A^T y with the fused primal update.
template <class Epi>
void launch_primal_half(pm_ctx* c, const Epi& ep, hipStream_t s) {
    if (c->merge_at)
        launch_merge(c->n, c->nnz, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(), ep, c->mpart_at.get(),
                     c->mlong_at.get(), c->nlong_at, c->mcarry.get(), c->mhead.get(), s);
    else if (c->sell_at.slices > 0) launch_sell_mat(c->sell_at, c->y.get(), ep, s);
    else launch_rows<float>(c->var_at, c->n, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(), ep, s);
}

// dst[q] = src[perm[q]], or 0 where perm[q] < 0 (padding).
__global__ __launch_bounds__(kBlock) void k_gather_pad_f32(int64_t n, const int* __restrict__ perm,
                                                           const float* __restrict__ src, float* __restrict__ dst) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (q < n) {
        const int k = perm[q];
        dst[q] = k >= 0 ? src[k] : 0.0f;
    }
}

// Builds the sliced ELL copy S of a matrix of nrows rows held as a host CSR
// structure (rp, ci) whose FP32 values lie on the device in CSR order (dv).
// Returns false (S left empty) when the padded entries would not fit int32
// offsets.
bool build_sell(pm_ctx* c, int nrows, const int* rp, const int* ci, const float* dv, SellMat& S) {
    S.reset();
    const int ns = (nrows + 63) / 64;
    std::vector<int> order(nrows);
    std::iota(order.begin(), order.end(), 0);
    auto rlen = [&](int r) { return rp[r + 1] - rp[r]; };
    for (int w0 = 0; w0 < nrows; w0 += kSellWindow) {
        const int w1 = std::min(nrows, w0 + kSellWindow);
        std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int a, int b) { return rlen(a) > rlen(b); });
    }
    std::vector<int> len(ns, 0), row(static_cast<size_t>(ns) * 64, -1), ptr(ns);
    int64_t total = 0;
    for (int sl = 0; sl < ns; ++sl) {
        int L = 0;
        for (int l = 0; l < 64; ++l) {
            const int q = sl * 64 + l;
            if (q < nrows) {
                row[q] = order[q];
                L = std::max(L, rlen(order[q]));
            }
        }
        len[sl] = L;
        ptr[sl] = static_cast<int>(std::min<int64_t>(total, INT32_MAX));
        total += int64_t(64) * L;
    }
    if (total >= INT32_MAX) return false;
    std::vector<int> col(std::max<int64_t>(total, 1), 0), perm(std::max<int64_t>(total, 1), -1);
    for (int sl = 0; sl < ns; ++sl)
        for (int l = 0; l < 64; ++l) {
            const int r = row[sl * 64 + l];
            const int rl = r >= 0 ? rlen(r) : 0;
            const int pad_col = rl > 0 ? ci[rp[r + 1] - 1] : 0;
            for (int j = 0; j < len[sl]; ++j) {
                const int64_t q = ptr[sl] + int64_t(j) * 64 + l;
                if (j < rl) {
                    col[q] = ci[rp[r] + j];
                    perm[q] = rp[r] + j;
                } else {
                    col[q] = pad_col;
                }
            }
        }
    hipStream_t s = c->stream;
    S.ptr.upload(ptr, s);
    S.len.upload(len, s);
    S.row.upload(row, s);
    S.col.upload(col, s);
    DevBuf<int> d_perm;
    d_perm.upload(perm, s);
    S.val.resize(col.size());
    hipLaunchKernelGGL(k_gather_pad_f32, dim3((col.size() + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       static_cast<int64_t>(col.size()), d_perm.get(), dv, S.val.get());
    PM_LAUNCHED();
    sync(c);
    S.slices = ns;
    return true;
}

// Times a sliced ELL product y = S x (mean of 5 after 2 warm-ups).
// Mean time of 5 launches of `go` after warm-ups.  A candidate whose first
// timed launch already takes more than 4x `bound` (the time to beat) is
// dropped with that figure: it cannot win, and on skewed rows a losing
// layout can take milliseconds a launch.
template <class F>
float time_launches(pm_ctx* c, const F& go, float bound) {
    float ms = 0.0f;
    go();
    PM_HIP(hipEventRecord(c->ev0, c->stream));
    go();
    PM_HIP(hipEventRecord(c->ev1, c->stream));
    PM_HIP(hipEventSynchronize(c->ev1));
    PM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms > 4.0f * bound) return ms;
    PM_HIP(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < 5; ++r) go();
    PM_HIP(hipEventRecord(c->ev1, c->stream));
    PM_HIP(hipEventSynchronize(c->ev1));
    PM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    return ms / 5.0f;
}

float time_sell(pm_ctx* c, const SellMat& S, const float* x, float* y, float bound = 1e30f) {
    EpiStore<float> epi{y};
    return time_launches(c, [&] { launch_sell_mat(S, x, epi, c->stream); }, bound);
}

float time_merge(pm_ctx* c, int nrows, const int* rp, const int* ci, const float* v, const float* x,
                 const int* part, const int4* longs, int nlong, float* y, float bound = 1e30f) {
    EpiStore<float> epi{y};
    return time_launches(
        c,
        [&] {
            launch_merge(nrows, c->nnz, rp, ci, v, x, epi, part, longs, nlong, c->mcarry.get(), c->mhead.get(),
                         c->stream);
        },
        bound);
}

// Builds the column-blocked copy of the scaled A with nblk blocks (the
// values gathered on the device from va through a permutation).
void build_blocks(pm_ctx* c, const pm_lp* lp, int nblk) {
    const int m = c->m, n = c->n, nnz = c->nnz;
    const int* rp = lp->row_ptr;
    const int* ci = lp->col_idx;
    std::vector<int> rpb(static_cast<size_t>(nblk) * (m + 1)), cib(nnz), perm(nnz);
    std::vector<int64_t> cnt(nblk, 0);
    auto block_of = [&](int j) { return static_cast<int>(static_cast<int64_t>(j) * nblk / n); };
    for (int k = 0; k < nnz; ++k) ++cnt[block_of(ci[k])];
    std::vector<int64_t> base(nblk + 1, 0);
    for (int b = 0; b < nblk; ++b) base[b + 1] = base[b] + cnt[b];
    std::vector<int64_t> cur(base.begin(), base.end() - 1);
    for (int i = 0; i < m; ++i) {
        for (int b = 0; b < nblk; ++b) rpb[static_cast<size_t>(b) * (m + 1) + i] = static_cast<int>(cur[b]);
        for (int k = rp[i]; k < rp[i + 1]; ++k) {  // columns ascending: blocks in order
            const int64_t q = cur[block_of(ci[k])]++;
            cib[q] = ci[k];
            perm[q] = k;
        }
    }
    for (int b = 0; b < nblk; ++b) rpb[static_cast<size_t>(b) * (m + 1) + m] = static_cast<int>(cur[b]);
    hipStream_t s = c->stream;
    c->rpb.upload(rpb, s);
    c->cib.upload(cib, s);
    DevBuf<int> d_perm;
    d_perm.upload(perm, s);
    c->vb.resize(std::max(nnz, 1));
    c->accb.resize(std::max(m, 1));
    hipLaunchKernelGGL(k_gather_f32, dim3((nnz + kBlock - 1) / kBlock), dim3(kBlock), 0, s, static_cast<int64_t>(nnz),
                       d_perm.get(), c->va.get(), c->vb.get());
    PM_LAUNCHED();
    sync(c);
    c->nblk = nblk;
}

// Times the column-blocked A x (mean of 5 after 2 warm-ups).
float time_blocked(pm_ctx* c, int nblk, int var, const float* x, float* y, float bound = 1e30f) {
    EpiStore<float> epi{y};
    return time_launches(c, [&] { launch_blocked(c, nblk, var, x, epi, c->stream); }, bound);
}

// Times one row-kernel SpMV variant.
float time_variant(pm_ctx* c, int variant, int nrows, const int* rp, const int* ci, const float* v,
                   const float* x, float* y, float bound = 1e30f) {
    EpiStore<float> epi{y};
    return time_launches(c, [&] { launch_rows<float>(variant, nrows, rp, ci, v, x, epi, c->stream); }, bound);
}

int autotune(pm_ctx* c, int nrows, int ncols, int nnz, const int* rp, const int* ci, const float* v,
             const float* x, float* y) {
    const int ov = c->opts.spmv_variant;
    if (ov != PM_SPMV_AUTO && ov != PM_SPMV_COLBLOCK && ov != PM_SPMV_SLICED && ov != PM_SPMV_MERGE) return ov;
    if (nnz < (1 << 18) || nrows == 0) return heuristic_variant(nrows, nnz);
    const auto key = std::make_tuple(nrows, ncols, nnz);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) return it->second;
    }
    const double avg = static_cast<double>(nnz) / nrows;
    int best = heuristic_variant(nrows, nnz);
    float best_ms = 1e30f;
    const int cands[] = {PM_SPMV_SCALAR,   PM_SPMV_VECTOR2,  PM_SPMV_VECTOR4, PM_SPMV_VECTOR8,
                         PM_SPMV_VECTOR16, PM_SPMV_VECTOR32, PM_SPMV_VECTOR64};
    for (int v_ : cands) {
        const int lanes = variant_lanes(v_);
        if (lanes > 2 * avg + 2 && lanes > 2) continue;  // mostly idle lanes
        const float ms = time_variant(c, v_, nrows, rp, ci, v, x, y, best_ms);
        if (ms < best_ms) { best_ms = ms; best = v_; }
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[key] = best;
    return best;
}

// ---------------------------------------------------------------------------
// Load
// ---------------------------------------------------------------------------
// Scales the LP whose original values are in vs (CSR) and vst (CSC), both
// overwritten with the scaled values; dr, dc receive the factors.
void device_ruiz_pc(pm_ctx* c, int passes, DevBuf<double>& vs, DevBuf<double>& vst, DevBuf<double>& dr,
                    DevBuf<double>& dc) {
    const int m = c->m, n = c->n;
    DevBuf<double> rn(std::max(m, 1)), cn(std::max(n, 1));
    dr.resize(std::max(m, 1));
    dc.resize(n);
    ruiz_pc_device(c->stream, passes, m, n, c->nnz, c->rp.get(), c->ci.get(), c->cp.get(), c->ri.get(), vs.get(),
                   vst.get(), dr.get(), dc.get(), rn.get(), cn.get());
    PM_LAUNCHED();
}

// Host-side checks of the options (no device work).
void check_options(const pm_pdhg_options& opts) {
    if (opts.check_frequency <= 0) fail(PM_ERR_INVALID_ARG, "check_frequency must be positive");
    if (opts.max_iterations < 0) fail(PM_ERR_INVALID_ARG, "max_iterations must be non-negative");
    if (opts.spmv_variant < 0 || opts.spmv_variant >= PM_SPMV_NUM_VARIANTS)
        fail(PM_ERR_INVALID_ARG, "unknown spmv_variant");
    if (opts.spmv_variant == PM_SPMV_STREAM || opts.spmv_variant == PM_SPMV_STAGED)
        fail(PM_ERR_INVALID_ARG, "spmv_variant PM_SPMV_STREAM and PM_SPMV_STAGED are batch-engine kernels");
    if ((opts.spmv_variant == PM_SPMV_SLICED || opts.spmv_variant == PM_SPMV_COLBLOCK ||
         opts.spmv_variant == PM_SPMV_MERGE) &&
        opts.algorithm != PM_ALG_PDLP)
        fail(PM_ERR_INVALID_ARG, "spmv_variant PM_SPMV_SLICED, PM_SPMV_COLBLOCK and PM_SPMV_MERGE need PM_ALG_PDLP");
    if (opts.algorithm != PM_ALG_REFERENCE && opts.algorithm != PM_ALG_PDLP)
        fail(PM_ERR_INVALID_ARG, "unknown algorithm");
    if (opts.termination != PM_TERM_REFERENCE && opts.termination != PM_TERM_RELATIVE_KKT)
        fail(PM_ERR_INVALID_ARG, "unknown termination");
    if (opts.restart_metric != PM_RESTART_KKT && opts.restart_metric != PM_RESTART_GAP)
        fail(PM_ERR_INVALID_ARG, "unknown restart metric");
    if (opts.step_rule != PM_STEP_HALPERN && opts.step_rule != PM_STEP_ADAPTIVE)
        fail(PM_ERR_INVALID_ARG, "unknown step rule");
    // The reference algorithm is the reference's solver unchanged: its own
    // test and its own starting point (pdhg_solver.cu:288-297,356).
    if (opts.algorithm == PM_ALG_REFERENCE && opts.termination != PM_TERM_REFERENCE)
        fail(PM_ERR_INVALID_ARG, "PM_ALG_REFERENCE stops on the reference test only; relative KKT needs PM_ALG_PDLP");
    if (opts.algorithm == PM_ALG_REFERENCE && (opts.x0 || opts.y0))
        fail(PM_ERR_INVALID_ARG, "PM_ALG_REFERENCE starts from x = clamp(0, l, u), y = 0; warm starts need PM_ALG_PDLP");
}

// Longest chunk captured as a graph (two or more launches per iteration).
constexpr int kGraphMaxChunk = 512;

// Drops the captured iteration graphs (hal_chunk, ref_chunk).
void drop_chunk_graph(pm_ctx* c) {
    if (c->chunk_exec) (void)hipGraphExecDestroy(c->chunk_exec);
    if (c->chunk_graph) (void)hipGraphDestroy(c->chunk_graph);
    if (c->ref_exec) (void)hipGraphExecDestroy(c->ref_exec);
    if (c->ref_graph) (void)hipGraphDestroy(c->ref_graph);
    c->chunk_exec = c->ref_exec = nullptr;
    c->chunk_graph = c->ref_graph = nullptr;
    c->chunk_len = c->ref_len = 0;
}

void ctx_load(pm_ctx* c, const pm_lp* lp, const pm_pdhg_options* o) {
    check_lp(lp);
    pm_pdhg_options opts;
    pm_pdhg_options_default(&opts);
    if (o) opts = *o;
    check_options(opts);
    PM_HIP(hipStreamSynchronize(c->stream));
    drop_chunk_graph(c);  // the layouts and buffers it names are rebuilt
    c->opts = opts;
    c->loaded = false;
    const int m = lp->m, n = lp->n, nnz = lp->nnz;
    c->m = m; c->n = n; c->nnz = nnz;
    c->sense = lp->sense;
    c->offset = lp->offset;
    hipStream_t s = c->stream;

    c->c_model.assign(lp->c, lp->c + n);
    c->nan_cost = false;
    for (int j = 0; j < n; ++j) if (!std::isfinite(lp->c[j])) c->nan_cost = true;
    const double sgn = lp->sense == 1 ? -1.0 : 1.0;
    c->cmin.resize(n);
    for (int j = 0; j < n; ++j) c->cmin[j] = sgn * lp->c[j];
    c->lr.assign(lp->row_lower, lp->row_lower + m);
    c->ur.assign(lp->row_upper, lp->row_upper + m);
    c->lx.assign(lp->var_lower, lp->var_lower + n);
    c->ux.assign(lp->var_upper, lp->var_upper + n);
    c->x0_h.clear(); c->y0_h.clear();
    if (opts.x0) c->x0_h.assign(opts.x0, opts.x0 + n);
    if (opts.y0) c->y0_h.assign(opts.y0, opts.y0 + m);
    c->opts.x0 = nullptr;
    c->opts.y0 = nullptr;
    if (n == 0 || c->nan_cost) { c->loaded = true; return; }

    for (auto* b : {&c->oa32, &c->oat32, &c->va, &c->vat}) b->reset();
    for (auto* b : {&c->oa64, &c->oat64, &c->d_dr, &c->d_dc}) b->reset();
    c->exact_f32 = true;
    for (int k = 0; k < nnz; ++k)
        if (static_cast<double>(static_cast<float>(lp->values[k])) != lp->values[k]) { c->exact_f32 = false; break; }

    std::vector<int> cp, ri;
    std::vector<double> vt64;
    host_transpose<double>(m, n, lp->row_ptr, lp->col_idx, lp->values, cp, ri, vt64, nullptr);
    c->rp.upload(lp->row_ptr, static_cast<size_t>(m) + 1, s);
    c->ci.upload(lp->col_idx, nnz, s);
    c->cp.upload(cp, s);
    c->ri.upload(ri, s);

    auto up64 = [&](DevBuf<double>& d, const std::vector<double>& h) { d.upload(h, s); };
    up64(c->d_cmin, c->cmin);
    up64(c->d_lr, c->lr);
    up64(c->d_ur, c->ur);
    up64(c->d_lx, c->lx);
    up64(c->d_ux, c->ux);

    // FP32 copies of bounds with the +-1e30f sentinel (pdhg_solver.cu:198-220).
    std::vector<float> lr32(m), ur32(m), lx32(n), ux32(n);
    for (int i = 0; i < m; ++i) {
        lr32[i] = c->lr[i] <= -1e29 ? -kFltInf : static_cast<float>(c->lr[i]);
        ur32[i] = c->ur[i] >= 1e29 ? kFltInf : static_cast<float>(c->ur[i]);
    }
    for (int j = 0; j < n; ++j) {
        lx32[j] = c->lx[j] <= -1e29 ? -kFltInf : static_cast<float>(c->lx[j]);
        ux32[j] = c->ux[j] >= 1e29 ? kFltInf : static_cast<float>(c->ux[j]);
    }

    const size_t pn = partials_needed(std::max(m, n), 16);
    c->partials.resize(pn);
    c->red.resize(64);
    c->hred.resize(64);
    for (auto* b : {&c->x, &c->xe, &c->x0, &c->xt, &c->xl, &c->tmp_n}) b->resize(n);
    for (auto* b : {&c->y, &c->y0, &c->yt, &c->yl, &c->tmp_m}) b->resize(std::max(m, 1));
    for (auto* b : {&c->xtot, &c->xbase, &c->xbest}) b->resize(n);
    for (auto* b : {&c->ytot, &c->ybase, &c->act64, &c->ybest}) b->resize(std::max(m, 1));
    c->adaptive_step = opts.algorithm == PM_ALG_PDLP && opts.step_rule == PM_STEP_ADAPTIVE;
    if (c->adaptive_step) {
        c->xs.resize(n);
        c->ys.resize(std::max(m, 1));
        c->adx.resize(std::max(m, 1));
    }
    c->restart_on_gap = opts.algorithm == PM_ALG_PDLP && opts.restart_metric == PM_RESTART_GAP;
    if (c->restart_on_gap) {
        c->ndg_part.resize(static_cast<size_t>((std::max(m, n) + kBlock - 1) / kBlock + 1) * 2 * kNdgT + 64);
        c->ndg_red.resize(4 * kNdgT);
        c->ndg_h.resize(4 * kNdgT);
    }
    if (opts.algorithm == PM_ALG_PDLP && opts.feasibility_polishing) {
        c->xsave.resize(n);
        c->xpol.resize(n);
        c->ysave.resize(std::max(m, 1));
        c->hold32.resize(6 * (static_cast<size_t>(n) + std::max(m, 1)));
        c->hold64.resize(static_cast<size_t>(n) + std::max(m, 1));
    }

    if (opts.algorithm == PM_ALG_REFERENCE) {
        // Canonicalisation and diagonal steps as pdhg_solver.cu:196-250.
        const float obj_sign = lp->sense == 1 ? -1.0f : 1.0f;
        std::vector<float> c32(n), v32(nnz), vt32(nnz), sig(m, 0.95f), tau(n, 0.95f);
        for (int j = 0; j < n; ++j) c32[j] = static_cast<float>(static_cast<double>(obj_sign) * lp->c[j]);
        for (int k = 0; k < nnz; ++k) v32[k] = static_cast<float>(lp->values[k]);
        for (int k = 0; k < nnz; ++k) vt32[k] = static_cast<float>(vt64[k]);
        std::vector<double> rn(m, 0.0), cn(n, 0.0);
        for (int i = 0; i < m; ++i)
            for (int k = lp->row_ptr[i]; k < lp->row_ptr[i + 1]; ++k) {
                const double av = std::abs(static_cast<double>(v32[k]));
                rn[i] += av;
                cn[lp->col_idx[k]] += av;
            }
        const float alpha = 0.95f;
        for (int i = 0; i < m; ++i) if (rn[i] > 1e-12) sig[i] = alpha / static_cast<float>(rn[i]);
        for (int j = 0; j < n; ++j) if (cn[j] > 1e-12) tau[j] = alpha / static_cast<float>(cn[j]);
        c->c_inf32 = 1.0;
        for (int j = 0; j < n; ++j) c->c_inf32 = std::max(c->c_inf32, static_cast<double>(std::abs(c32[j])));
        c->va.upload(v32, s);
        c->vat.upload(vt32, s);
        if (!c->exact_f32) {
            c->oa64.upload(lp->values, nnz, s);
            c->oat64.upload(vt64, s);
        }
        up64(c->d_dr, std::vector<double>(std::max(m, 1), 1.0));
        up64(c->d_dc, std::vector<double>(n, 1.0));
        c->c32.upload(c32, s);
        c->sig32.upload(sig, s);
        c->tau32.upload(tau, s);
        c->lr32.upload(lr32, s);
        c->ur32.upload(ur32, s);
        c->lx32.upload(lx32, s);
        c->ux32.upload(ux32, s);
    } else {
        // Ruiz equilibration of A (inf-norm, opts.ruiz_iterations passes)
        // followed by Pock-Chambolle alpha = 1 (1-norms), on the device.
        DevBuf<double> vs, vst;
        vs.upload(lp->values, nnz, s);
        vst.upload(vt64, s);
        if (c->exact_f32) {
            std::vector<float> o(nnz), ot(nnz);
            for (int k = 0; k < nnz; ++k) o[k] = static_cast<float>(lp->values[k]);
            for (int k = 0; k < nnz; ++k) ot[k] = static_cast<float>(vt64[k]);
            c->oa32.upload(o, s);
            c->oat32.upload(ot, s);
        } else {
            c->oa64.resize(nnz);
            c->oat64.resize(nnz);
            PM_HIP(hipMemcpyAsync(c->oa64.get(), vs.get(), sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
            PM_HIP(hipMemcpyAsync(c->oat64.get(), vst.get(), sizeof(double) * nnz, hipMemcpyDeviceToDevice, s));
        }
        device_ruiz_pc(c, opts.ruiz_iterations, vs, vst, c->d_dr, c->d_dc);
        c->va.resize(std::max(nnz, 1));
        c->vat.resize(std::max(nnz, 1));
        if (nnz > 0) {
            const dim3 gz((nnz + kBlock - 1) / kBlock), b(kBlock);
            hipLaunchKernelGGL(k_to_f32, gz, b, 0, s, static_cast<int64_t>(nnz), vs.get(), c->va.get());
            hipLaunchKernelGGL(k_to_f32, gz, b, 0, s, static_cast<int64_t>(nnz), vst.get(), c->vat.get());
            PM_LAUNCHED();
        }
        std::vector<double> dr(std::max(m, 1)), dc(n);
        if (m > 0) c->d_dr.download(dr.data(), m, s);
        c->d_dc.download(dc.data(), n, s);
        sync(c);
        for (auto* b : {&c->c32, &c->lx32, &c->ux32}) b->resize(n);
        for (auto* b : {&c->lr32, &c->ur32, &c->g32}) b->resize(std::max(m, 1));

        // Initial primal weight ||c_hat||_2 / ||b_hat||_2 (PDLP).
        double cn2 = 0.0, bn2 = 0.0;
        for (int j = 0; j < n; ++j) cn2 += (dc[j] * c->cmin[j]) * (dc[j] * c->cmin[j]);
        for (int i = 0; i < m; ++i) {
            double bh = 0.0;
            if (c->lr[i] > -kInf) bh = std::max(bh, std::abs(c->lr[i]));
            if (c->ur[i] < kInf) bh = std::max(bh, std::abs(c->ur[i]));
            bn2 += (dr[i] * bh) * (dr[i] * bh);
        }
        c->omega0 = (cn2 > 1e-20 && bn2 > 1e-20) ? std::sqrt(cn2) / std::sqrt(bn2) : 1.0;
        if (opts.primal_weight > 0.0) c->omega0 = opts.primal_weight;
    }

    // SpMV variant per orientation.
    {
        Range range_("pm_ctx_load: kernel choice");
        float* xs = c->x.get();
        launch_elem(n, FnFill<float>{xs, 1.0f}, s);
        launch_elem(m, FnFill<float>{c->y.get(), 1.0f}, s);
        c->var_a = autotune(c, m, n, nnz, c->rp.get(), c->ci.get(), c->va.get(), xs, c->tmp_m.get());
        c->var_at = autotune(c, n, m, nnz, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(), c->tmp_n.get());
        // Sliced ELL copies for the iteration's products (PDLP): built and
        // timed beside the row kernels on LPs of at least 2^18 nonzeros,
        // kept where faster (always under PM_SPMV_SLICED).
        c->sell_a.reset();
        c->sell_at.reset();
        float t_a = -1.0f;  // time of A x-bar by the kernel kept, when measured
        const bool want_sell = opts.spmv_variant == PM_SPMV_SLICED;
        if (opts.algorithm == PM_ALG_PDLP && m > 0 && nnz > 0 &&
            (want_sell || ((opts.spmv_variant == PM_SPMV_AUTO || opts.spmv_variant == PM_SPMV_COLBLOCK) &&
                           nnz >= (1 << 18)))) {
            if (build_sell(c, n, cp.data(), ri.data(), c->vat.get(), c->sell_at) && !want_sell) {
                const float tr = time_variant(c, c->var_at, n, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(),
                                              c->tmp_n.get());
                if (time_sell(c, c->sell_at, c->y.get(), c->tmp_n.get(), tr) >= tr) c->sell_at.reset();
            }
            if (build_sell(c, m, lp->row_ptr, lp->col_idx, c->va.get(), c->sell_a) && !want_sell) {
                const float tr = time_variant(c, c->var_a, m, c->rp.get(), c->ci.get(), c->va.get(), xs,
                                              c->tmp_m.get());
                const float ts = time_sell(c, c->sell_a, xs, c->tmp_m.get(), tr);
                t_a = std::min(tr, ts);
                if (ts >= tr) c->sell_a.reset();
            }
        }
        // Merge-path products: timed against what was kept for each product
        // (or forced by PM_SPMV_MERGE); they even out skewed row lengths.
        c->merge_a = c->merge_at = false;
        const bool want_merge = opts.spmv_variant == PM_SPMV_MERGE;
        if (opts.algorithm == PM_ALG_PDLP && m > 0 && nnz > 0 &&
            (want_merge || (opts.spmv_variant == PM_SPMV_AUTO && nnz >= (1 << 18)))) {
            const size_t tiles = static_cast<size_t>(std::max(merge_tiles(m, nnz), merge_tiles(n, nnz)));
            c->mcarry.resize(tiles);
            c->mhead.resize(tiles);
            c->mpart_a.resize(static_cast<size_t>(merge_tiles(m, nnz)) + 1);
            c->mpart_at.resize(static_cast<size_t>(merge_tiles(n, nnz)) + 1);
            launch_merge_partition(m, nnz, c->rp.get(), c->mpart_a.get(), s);
            launch_merge_partition(n, nnz, c->cp.get(), c->mpart_at.get(), s);
            const std::vector<int4> la = merge_long_rows(m, lp->row_ptr), lat = merge_long_rows(n, cp.data());
            c->nlong_a = static_cast<int>(la.size());
            c->nlong_at = static_cast<int>(lat.size());
            c->mlong_a.upload(la, s);
            c->mlong_at.upload(lat, s);
            const float tc_at = want_merge ? 1e30f
                                : c->sell_at.slices > 0 ? time_sell(c, c->sell_at, c->y.get(), c->tmp_n.get())
                                                        : time_variant(c, c->var_at, n, c->cp.get(), c->ri.get(),
                                                                       c->vat.get(), c->y.get(), c->tmp_n.get());
            const float tm_at = time_merge(c, n, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(),
                                           c->mpart_at.get(), c->mlong_at.get(), c->nlong_at, c->tmp_n.get(), tc_at);
            if (tm_at < tc_at) {
                c->merge_at = true;
                c->sell_at.reset();
            }
            const float tc_a = want_merge ? 1e30f
                               : t_a >= 0.0f ? t_a
                                             : time_variant(c, c->var_a, m, c->rp.get(), c->ci.get(), c->va.get(), xs,
                                                            c->tmp_m.get());
            const float tm_a = time_merge(c, m, c->rp.get(), c->ci.get(), c->va.get(), xs, c->mpart_a.get(),
                                          c->mlong_a.get(), c->nlong_a, c->tmp_m.get(), tc_a);
            if (tm_a < tc_a) {
                c->merge_a = true;
                c->sell_a.reset();
                t_a = tm_a;
            }
            if (!c->merge_a) c->mpart_a.reset();
            if (!c->merge_at) c->mpart_at.reset();
            if (!c->merge_a && !c->merge_at) {
                c->mcarry.reset();
                c->mhead.reset();
            }
        }
        // Column blocks for the iteration's A x-bar: tried on large LPs,
        // where x-bar outgrows what an XCD's L2 keeps beside the matrix
        // stream (the 500k x 1M LP fetched 4.2x its algorithmic bytes in
        // plain A x-bar, most of it lines brought in for one gathered value).
        c->nblk = 0;
        const bool forced = opts.spmv_variant == PM_SPMV_COLBLOCK;
        if (opts.algorithm == PM_ALG_PDLP && m > 0 && nnz > 0 &&
            (forced || (opts.spmv_variant == PM_SPMV_AUTO && n >= (1 << 19) && nnz >= (1 << 22)))) {
            const int slab = std::max(2, static_cast<int>((static_cast<int64_t>(n) * 4 + (1 << 20) - 1) >> 20));
            float best = forced ? 1e30f
                         : t_a >= 0.0f ? t_a
                                       : time_variant(c, c->var_a, m, c->rp.get(), c->ci.get(), c->va.get(), xs,
                                                      c->tmp_m.get());
            int best_nb = 0, best_var = PM_SPMV_VECTOR4;
            std::vector<int> counts{2, slab, 2 * slab};  // x-bar in halves, in slabs <= 1 MB, <= 512 KB
            counts.erase(std::unique(counts.begin(), counts.end()), counts.end());
            for (int nb : counts) {
                build_blocks(c, lp, nb);
                for (int v_ : {PM_SPMV_VECTOR2, PM_SPMV_VECTOR4, PM_SPMV_VECTOR8}) {
                    const float ms = time_blocked(c, nb, v_, xs, c->tmp_m.get(), best);
                    if (ms < best) { best = ms; best_nb = nb; best_var = v_; }
                }
            }
            if (best_nb > 0) {
                if (best_nb != c->nblk) build_blocks(c, lp, best_nb);
                c->blk_var = best_var;
                c->sell_a.reset();
                c->merge_a = false;
                c->mpart_a.reset();
            } else {
                c->nblk = 0;
                for (auto* bf : {&c->rpb, &c->cib}) bf->reset();
                c->vb.reset();
                c->accb.reset();
            }
        }
    }

    if (opts.algorithm == PM_ALG_PDLP && m > 0 && nnz > 0) {
        Range range_("pm_ctx_load: power iteration");
        // Power iteration for ||A_hat||_2 on A^T A.  Its estimate approaches
        // sigma_max from below, so it runs until the estimate settles (in
        // blocks of 16 steps after the first 64, relative change of
        // ||A^T A v||^2 <= 1e-5, at most 1,024 steps): with clustered top
        // singular values a fixed count can stop a percent short, and
        // eta = 0.99 / sigma_est would then leave little of the margin.
        float* v = c->x.get();
        float* w = c->tmp_n.get();
        float* u = c->tmp_m.get();
        double* ss = c->red.get();
        launch_elem(n, FnFill<float>{v, 1.0f / std::sqrt(static_cast<float>(n))}, s);
        double lam = 0.0, prev = -1.0;
        for (int it = 0; it < 1024;) {
            const int steps = it == 0 ? 64 : 16;
            for (int q = 0; q < steps; ++q) {
                launch_rows<float>(c->var_a, m, c->rp.get(), c->ci.get(), c->va.get(), v, EpiStore<float>{u}, s);
                launch_rows<float>(c->var_at, n, c->cp.get(), c->ri.get(), c->vat.get(), u, EpiStore<float>{w}, s);
                launch_elem_reduce(n, FnNormSq{w}, c->partials.get(), ss, s);
                launch_elem(n, FnNormalize{w, ss, v}, s);
            }
            it += steps;
            PM_HIP(hipMemcpyAsync(c->hred.get(), ss, sizeof(double), hipMemcpyDeviceToHost, s));
            sync(c);
            lam = c->hred[0];  // ||A^T A v||^2 with ||v|| = 1 -> sigma_max^4
            if (!(lam > 0.0) || std::fabs(lam - prev) <= 1e-5 * lam) break;
            prev = lam;
        }
        c->sigma_max = lam > 0.0 ? std::sqrt(std::sqrt(lam)) : 1.0;
        c->eta = 0.99 / c->sigma_max;
    }
    sync(c);
    c->loaded = true;
}

// ---------------------------------------------------------------------------
// Reference solve
// ---------------------------------------------------------------------------
void ref_residuals(pm_ctx* c, double tol, double& pv, double& dv) {
    hipStream_t s = c->stream;
    double* red = c->red.get();
    const int m = c->m, n = c->n;
    const int vrow = c->opts.spmv_variant == PM_SPMV_SEQUENTIAL ? PM_SPMV_SEQUENTIAL : c->var_a;
    const int vcol = c->opts.spmv_variant == PM_SPMV_SEQUENTIAL ? PM_SPMV_SEQUENTIAL : c->var_at;
    EpiRefRowViol er{c->d_lr.get(), c->d_ur.get()};
    if (c->exact_f32)
        launch_rows_reduce<double>(vrow, m, c->rp.get(), c->ci.get(), c->va.get(), c->x.get(), er,
                                   c->partials.get(), red, s);
    else
        launch_rows_reduce<double>(vrow, m, c->rp.get(), c->ci.get(), c->oa64.get(), c->x.get(), er,
                                   c->partials.get(), red, s);
    EpiRefColDual<float> ec{c->x.get(), c->c32.get(), c->d_lx.get(), c->d_ux.get(), tol};
    launch_rows_reduce<double>(vcol, n, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(), ec,
                               c->partials.get(), red + 1, s);
    PM_HIP(hipMemcpyAsync(c->hred.get(), red, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    pv = std::max(c->hred[0], c->hred[1]);
    dv = c->hred[2] / c->c_inf32;
}

void ref_iteration(pm_ctx* c) {
    hipStream_t s = c->stream;
    const int va = c->var_a, vt = c->var_at;
    if (c->m > 0)
        launch_rows<float>(va, c->m, c->rp.get(), c->ci.get(), c->va.get(), c->xe.get(),
                           EpiDualRef{c->sig32.get(), c->lr32.get(), c->ur32.get(), c->y.get()}, s);
    launch_rows<float>(vt, c->n, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(),
                       EpiPrimalRef{c->tau32.get(), c->c32.get(), c->lx32.get(), c->ux32.get(), c->x.get(),
                                    c->xe.get()},
                       s);
}

// `len` reference iterations; on an LP below 2^18 nonzeros a run of K
// (check_frequency) replays one captured graph of the same launches.
void ref_chunk(pm_ctx* c, int len, int K) {
    if (!c->use_graph || len != K || K > kGraphMaxChunk || c->nnz >= (1 << 18)) {
        for (int i = 0; i < len; ++i) ref_iteration(c);
        return;
    }
    hipStream_t s = c->stream;
    if (!c->ref_exec || c->ref_len != K) {
        if (c->ref_exec) (void)hipGraphExecDestroy(c->ref_exec);
        if (c->ref_graph) (void)hipGraphDestroy(c->ref_graph);
        c->ref_exec = nullptr;
        c->ref_graph = nullptr;
        PM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < K; ++i) ref_iteration(c);
        PM_HIP(hipStreamEndCapture(s, &c->ref_graph));
        PM_HIP(hipGraphInstantiate(&c->ref_exec, c->ref_graph, nullptr, nullptr, 0));
        c->ref_len = K;
    }
    PM_HIP(hipGraphLaunch(c->ref_exec, s));
}

void finish_result(pm_ctx* c, pm_pdhg_result* res, const double* xd_dev, const double* yd_dev,
                   std::vector<double>& xh, std::vector<double>& yh);

void ref_solve(pm_ctx* c, pm_pdhg_result* res) {
    hipStream_t s = c->stream;
    const int m = c->m, n = c->n;
    const auto& o = c->opts;
    if (c->var_a == PM_SPMV_SEQUENTIAL || o.spmv_variant == PM_SPMV_SEQUENTIAL) {
        c->var_a = PM_SPMV_SEQUENTIAL;
        c->var_at = PM_SPMV_SEQUENTIAL;
    }
    // x0 = clamp(0, l, u), x-bar = x0, y = 0 (pdhg_solver.cu:288-297).
    launch_elem(n, FnInitX{c->lx32.get(), c->ux32.get(), c->x.get(), c->xe.get(), nullptr}, s);
    launch_elem(m, FnFill<float>{c->y.get(), 0.0f}, s);
    double best = INFINITY;
    int stagnation = 0, restarts = 0;
    const int kRestP = 8, kMaxR = 10;
    res->status = PM_SOLVE_ITERATION_LIMIT;
    const auto t0 = std::chrono::steady_clock::now();
    const int K = o.check_frequency;
    for (int done = 0; done < o.max_iterations;) {
        // iterations up to the next check (every K-th iteration), or the end
        const int next = std::min(o.max_iterations, (done / K + 1) * K);
        ref_chunk(c, next - done, K);
        done = next;
        if (done % K == 0) {
            double pv = 0, dv = 0;
            ref_residuals(c, o.tolerance, pv, dv);
            res->primal_residual = pv;
            res->dual_residual = dv;
            if (o.verbose)
                fprintf(stderr, "[pramaan] iter %d pv %.3e dv %.3e restarts %d\n", done, pv, dv, restarts);
            if (pv <= o.tolerance && dv <= o.tolerance) {
                res->status = PM_SOLVE_OPTIMAL;
                res->iterations = done;
                break;
            }
            const double combined = pv + dv;
            if (combined > o.tolerance * 10.0) {
                if (combined < best * 0.99) { best = combined; stagnation = 0; }
                else ++stagnation;
                if (stagnation >= kRestP && restarts < kMaxR) {
                    launch_elem(m, FnFill<float>{c->y.get(), 0.0f}, s);
                    launch_elem(n, FnCopy<float>{c->xe.get(), c->x.get()}, s);
                    stagnation = 0;
                    best = INFINITY;
                    ++restarts;
                }
            }
        }
    }
    sync(c);
    res->device_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (res->status != PM_SOLVE_OPTIMAL) res->iterations = o.max_iterations;
    double pv = 0, dv = 0;
    ref_residuals(c, o.tolerance, pv, dv);
    res->primal_residual = pv;
    res->dual_residual = dv;
    res->restarts = restarts;
    // FP64 copies of the FP32 iterates for the KKT evaluation and outputs.
    std::vector<float> xf(n), yf(m);
    c->x.download(xf.data(), n, s);
    if (m > 0) c->y.download(yf.data(), m, s);
    sync(c);
    std::vector<double> xh(n), yh(m);
    for (int j = 0; j < n; ++j) xh[j] = static_cast<double>(xf[j]);
    for (int i = 0; i < m; ++i) yh[i] = static_cast<double>(yf[i]);
    c->xtot.upload(xh, s);
    if (m > 0) c->ytot.upload(yh, s);
    finish_result(c, res, c->xtot.get(), c->ytot.get(), xh, yh);
}

// Relative KKT of (xtot, ytot) on the device.  out: see EpiKktRow/Col.
struct Kkt {
    double rel, rel_p, rel_d, rel_g, p, d, rp_inf, rd_inf, ps, ds;
};

Kkt eval_kkt(pm_ctx* c, const double* xd, const double* yd) {
    hipStream_t s = c->stream;
    double* red = c->red.get();
    const int m = c->m, n = c->n;
    const int vrow = c->opts.spmv_variant == PM_SPMV_SEQUENTIAL ? PM_SPMV_SEQUENTIAL : c->var_a;
    const int vcol = c->opts.spmv_variant == PM_SPMV_SEQUENTIAL ? PM_SPMV_SEQUENTIAL : c->var_at;
    const double* dr = c->d_dr.get();   // unit scales in reference mode
    const double* dcp = c->d_dc.get();
    EpiKktRow er{c->d_lr.get(), c->d_ur.get(), yd, dr};
    EpiKktCol ec{c->d_cmin.get(), c->d_lx.get(), c->d_ux.get(), xd, dcp};
    const bool ref = c->opts.algorithm == PM_ALG_REFERENCE;
    if (ref && c->exact_f32) {
        launch_rows_reduce<double>(vrow, m, c->rp.get(), c->ci.get(), c->va.get(), xd, er, c->partials.get(), red, s);
        launch_rows_reduce<double>(vcol, n, c->cp.get(), c->ri.get(), c->vat.get(), yd, ec, c->partials.get(),
                                   red + 8, s);
    } else if (ref) {
        launch_rows_reduce<double>(vrow, m, c->rp.get(), c->ci.get(), c->oa64.get(), xd, er, c->partials.get(), red,
                                   s);
        launch_rows_reduce<double>(vcol, n, c->cp.get(), c->ri.get(), c->oat64.get(), yd, ec, c->partials.get(),
                                   red + 8, s);
    } else if (c->exact_f32) {
        launch_rows_reduce<double>(vrow, m, c->rp.get(), c->ci.get(), c->oa32.get(), xd, er, c->partials.get(), red,
                                   s);
        launch_rows_reduce<double>(vcol, n, c->cp.get(), c->ri.get(), c->oat32.get(), yd, ec, c->partials.get(),
                                   red + 8, s);
    } else {
        launch_rows_reduce<double>(vrow, m, c->rp.get(), c->ci.get(), c->oa64.get(), xd, er, c->partials.get(), red,
                                   s);
        launch_rows_reduce<double>(vcol, n, c->cp.get(), c->ri.get(), c->oat64.get(), yd, ec, c->partials.get(),
                                   red + 8, s);
    }
    PM_HIP(hipMemcpyAsync(c->hred.get(), red, 14 * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    const double* h = c->hred.get();
    // row: 0 r^2, 1 (Dr r)^2, 2 bhat^2, 3 e^2, 4 (e/Dr)^2, 5 d_row, 6 |r|max
    // col: 8 e^2, 9 (Dc e)^2, 10 c^2, 11 d_col, 12 c^T x, 13 |e|max
    Kkt k{};
    k.p = h[12];
    k.d = h[5] + h[11];
    k.rel_p = std::sqrt(h[0]) / (1.0 + std::sqrt(h[2]));
    k.rel_d = std::sqrt(h[3] + h[8]) / (1.0 + std::sqrt(h[10]));
    k.rel_g = std::fabs(k.p - k.d) / (1.0 + std::fabs(k.p) + std::fabs(k.d));
    k.rel = std::max(k.rel_p, std::max(k.rel_d, k.rel_g));
    if (!std::isfinite(k.rel)) k.rel = INFINITY;
    k.rp_inf = h[6];
    k.rd_inf = h[13];
    k.ps = std::sqrt(h[1]);
    k.ds = std::sqrt(h[4] + h[9]);
    return k;
}

// Reference-style pv/dv (pdhg_solver.cu:107-161) of FP64 iterates.
void ref_style_residuals(pm_ctx* c, const double* xd, const double* yd, double tol, double& pv, double& dv) {
    hipStream_t s = c->stream;
    double* red = c->red.get() + 32;
    const int m = c->m, n = c->n;
    EpiRefRowViol er{c->d_lr.get(), c->d_ur.get()};
    if (c->exact_f32)
        launch_rows_reduce<double>(c->var_a, m, c->rp.get(), c->ci.get(),
                                   c->oa32.get() ? c->oa32.get() : c->va.get(), xd, er, c->partials.get(), red, s);
    else
        launch_rows_reduce<double>(c->var_a, m, c->rp.get(), c->ci.get(), c->oa64.get(), xd, er, c->partials.get(),
                                   red, s);
    // c in FP32 as the reference keeps it.
    std::vector<float> c32(n);
    for (int j = 0; j < n; ++j) c32[j] = static_cast<float>(c->cmin[j]);
    c->tmp_n.upload(c32, s);
    EpiRefColDual<double> ec{xd, c->tmp_n.get(), c->d_lx.get(), c->d_ux.get(), tol};
    if (c->exact_f32)
        launch_rows_reduce<double>(c->var_at, n, c->cp.get(), c->ri.get(),
                                   c->oat32.get() ? c->oat32.get() : c->vat.get(), yd, ec, c->partials.get(), red + 1,
                                   s);
    else
        launch_rows_reduce<double>(c->var_at, n, c->cp.get(), c->ri.get(), c->oat64.get(), yd, ec,
                                   c->partials.get(), red + 1, s);
    PM_HIP(hipMemcpyAsync(c->hred.get() + 32, red, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    double cinf = 1.0;
    for (int j = 0; j < n; ++j) cinf = std::max(cinf, static_cast<double>(std::abs(c32[j])));
    pv = std::max(c->hred[32], c->hred[33]);
    dv = c->hred[34] / cinf;
}

void finish_result(pm_ctx* c, pm_pdhg_result* res, const double* xd_dev, const double* yd_dev,
                   std::vector<double>& xh, std::vector<double>& yh) {
    const int m = c->m, n = c->n;
    hipStream_t s = c->stream;
    // Objective in the model's sense with the offset, from x (pdhg_solver.cu:402-408).
    double obj = c->offset;
    for (int j = 0; j < n; ++j) obj += c->c_model[j] * xh[j];
    res->objective = obj;
    {
        const Kkt k = eval_kkt(c, xd_dev, yd_dev);
        res->rel_kkt = k.rel;
        res->rel_primal = k.rel_p;
        res->rel_dual = k.rel_d;
        res->rel_gap = k.rel_g;
        res->primal_objective = k.p;
        res->dual_objective = k.d;
    }
    if (res->x) std::copy(xh.begin(), xh.end(), res->x);
    if (res->y && m > 0) std::copy(yh.begin(), yh.end(), res->y);
    if (res->row_activity && m > 0) {
        double* act = c->act64.get();
        if (c->exact_f32) {
            const float* v = (c->opts.algorithm == PM_ALG_PDLP) ? c->oa32.get() : c->va.get();
            launch_rows<double>(c->var_a, m, c->rp.get(), c->ci.get(), v, xd_dev, EpiStore<double>{act}, s);
        } else {
            launch_rows<double>(c->var_a, m, c->rp.get(), c->ci.get(), c->oa64.get(), xd_dev, EpiStore<double>{act},
                                s);
        }
        c->act64.download(res->row_activity, m, s);
        sync(c);
    }
}

// ---------------------------------------------------------------------------
// PDLP (reflected Halpern) solve
// ---------------------------------------------------------------------------
void round_data(pm_ctx* c, double dp, double dd, int mode = kRoundLp) {
    hipStream_t s = c->stream;
    const int m = c->m, n = c->n;
    EpiRoundCol ec{c->d_cmin.get(), c->d_lx.get(), c->d_ux.get(), c->xbase.get(), c->d_dc.get(), dp, dd,
                   c->c32.get(), c->lx32.get(), c->ux32.get(), mode};
    if (c->exact_f32)
        launch_rows<double>(c->var_at, n, c->cp.get(), c->ri.get(), c->oat32.get(), c->ybase.get(), ec, s);
    else
        launch_rows<double>(c->var_at, n, c->cp.get(), c->ri.get(), c->oat64.get(), c->ybase.get(), ec, s);
    if (m > 0) {
        EpiRoundRow er{c->d_lr.get(), c->d_ur.get(), c->ybase.get(), c->d_dr.get(), dp, dd,
                       c->lr32.get(), c->ur32.get(), c->g32.get(), mode};
        if (c->exact_f32)
            launch_rows<double>(c->var_a, m, c->rp.get(), c->ci.get(), c->oa32.get(), c->xbase.get(), er, s);
        else
            launch_rows<double>(c->var_a, m, c->rp.get(), c->ci.get(), c->oa64.get(), c->xbase.get(), er, s);
    }
    // Iterates restart at x = clamp(0, l_hat, u_hat), y = 0.
    launch_elem(n, FnInitX{c->lx32.get(), c->ux32.get(), c->x.get(), c->x0.get(), c->xl.get()}, s);
    launch_elem(m, FnFill<float>{c->y.get(), 0.0f}, s);
    launch_elem(m, FnFill<float>{c->y0.get(), 0.0f}, s);
    launch_elem(m, FnFill<float>{c->yl.get(), 0.0f}, s);
}

void hal_iteration(pm_ctx* c, int k, double omega, bool keep_t, bool use_g) {
    hipStream_t s = c->stream;
    const float tau = static_cast<float>(c->eta / omega);
    const float sig = static_cast<float>(c->eta * omega);
    const float lam = static_cast<float>((k + 1.0) / (k + 2.0));
    const float rho = 1.0f;
    EpiPrimalHal ep{tau, lam, rho, c->c32.get(), c->lx32.get(), c->ux32.get(), c->x0.get(), c->x.get(), c->xe.get(),
                    keep_t ? c->xt.get() : nullptr};
    launch_primal_half(c, ep, s);
    if (c->m > 0) {
        EpiDualHal ed{sig, lam, rho, c->lr32.get(), c->ur32.get(), use_g ? c->g32.get() : nullptr, c->y0.get(),
                      c->y.get(), keep_t ? c->yt.get() : nullptr};
        launch_dual_half(c, ed, s);
    }
}



// `chunk` iterations from in-round index k0 (the last keeps T(z)).  On an
// LP below 2^18 nonzeros a full chunk of K iterations replays a captured
// graph whose epilogues
// (EpiPrimalHalDev, EpiDualHalDev) read tau, sigma, the dual shift's
// presence and each position's Halpern weight from chunk_dp, filled here
// with the values hal_iteration would pass, so both paths give the same
// bits; a shorter chunk launches directly.  The buffers the graph names
// keep their addresses until the next load, which drops it.
void hal_chunk(pm_ctx* c, int k0, double omega, bool use_g, int chunk, int K) {
    // Graphs pay where launches cost as much as the kernels: LPs below the
    // autotuner's 2^18 nonzeros (larger products hide the launches).
    if (!c->use_graph || chunk != K || K > kGraphMaxChunk || c->nnz >= (1 << 18)) {
        for (int i = 0; i < chunk; ++i) hal_iteration(c, k0 + i, omega, i == chunk - 1, use_g);
        return;
    }
    hipStream_t s = c->stream;
    if (c->chunk_dp.size() < static_cast<size_t>(K) + 3) {
        c->chunk_dp.resize(static_cast<size_t>(K) + 3);
        c->chunk_hp.resize(static_cast<size_t>(K) + 3);
    }
    // the previous chunk's copy has completed: every check synchronises
    float* hp = c->chunk_hp.get();
    hp[0] = static_cast<float>(c->eta / omega);
    hp[1] = static_cast<float>(c->eta * omega);
    hp[2] = use_g ? 1.0f : 0.0f;
    for (int i = 0; i < K; ++i) {
        const int k = k0 + i;
        hp[3 + i] = static_cast<float>((k + 1.0) / (k + 2.0));
    }
    PM_HIP(hipMemcpyAsync(c->chunk_dp.get(), hp, sizeof(float) * (static_cast<size_t>(K) + 3),
                          hipMemcpyHostToDevice, s));
    if (!c->chunk_exec || c->chunk_len != K) {
        drop_chunk_graph(c);
        PM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < K; ++i) {
            const bool keep = i == K - 1;
            launch_primal_half(c,
                               EpiPrimalHalDev{c->chunk_dp.get(), i, c->c32.get(), c->lx32.get(), c->ux32.get(),
                                               c->x0.get(), c->x.get(), c->xe.get(), keep ? c->xt.get() : nullptr},
                               s);
            if (c->m > 0)
                launch_dual_half(c,
                                 EpiDualHalDev{c->chunk_dp.get(), i, c->lr32.get(), c->ur32.get(), c->g32.get(),
                                               c->y0.get(), c->y.get(), keep ? c->yt.get() : nullptr},
                                 s);
        }
        PM_HIP(hipStreamEndCapture(s, &c->chunk_graph));
        PM_HIP(hipGraphInstantiate(&c->chunk_exec, c->chunk_graph, nullptr, nullptr, 0));
        c->chunk_len = K;
    }
    PM_HIP(hipGraphLaunch(c->chunk_exec, s));
}

// One step of plain PDHG with PDLP's adaptive step size (PM_STEP_ADAPTIVE):
// a step of size eta (primal eta / omega, dual eta omega) is kept when eta
// <= eta_max = ||dz||^2_omega / (2 |dy^T A dx|), the largest step the
// PDHG convergence condition allows for that move; either way the next
// size is min((1 - (k+1)^-0.3) eta_max, (1 + (k+1)^-0.6) eta), k >= 1 the
// steps taken so far, and a rejected step is retried from the saved iterate.
// The Halpern epilogues with lambda = 1 and no reflection are plain PDHG.
// Returns the tries (each costs an iteration).
int adaptive_iteration(pm_ctx* c, double omega, bool keep_t, bool use_g, int k) {
    hipStream_t s = c->stream;
    const int m = c->m, n = c->n;
    double* red = c->red.get();
    for (int tries = 1;; ++tries) {
        launch_elem(n, FnCopy<float>{c->xs.get(), c->x.get()}, s);
        launch_elem(m, FnCopy<float>{c->ys.get(), c->y.get()}, s);
        const double eta = c->eta_ad;
        EpiPrimalHal ep{static_cast<float>(eta / omega), 1.0f, 0.0f, c->c32.get(), c->lx32.get(), c->ux32.get(),
                        c->x0.get(), c->x.get(), c->xe.get(), keep_t ? c->xt.get() : nullptr};
        launch_primal_half(c, ep, s);
        if (m > 0) {
            EpiDualHal ed{static_cast<float>(eta * omega), 1.0f, 0.0f, c->lr32.get(), c->ur32.get(),
                          use_g ? c->g32.get() : nullptr, c->y0.get(), c->y.get(), keep_t ? c->yt.get() : nullptr};
            launch_dual_half(c, ed, s);
        }
        launch_elem(n, FnSubF{c->tmp_n.get(), c->x.get(), c->xs.get()}, s);
        if (m > 0)
            launch_rows<float>(c->var_a, m, c->rp.get(), c->ci.get(), c->va.get(), c->tmp_n.get(),
                               EpiStore<float>{c->adx.get()}, s);
        launch_elem_reduce(n, FnNormDiff{c->x.get(), c->xs.get()}, c->partials.get(), red + 56, s);
        launch_elem_reduce(m, FnNormDiff{c->y.get(), c->ys.get()}, c->partials.get(), red + 57, s);
        launch_elem_reduce(m, FnDyAdx{c->y.get(), c->ys.get(), c->adx.get()}, c->partials.get(), red + 58, s);
        PM_HIP(hipMemcpyAsync(c->hred.get() + 56, red + 56, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        sync(c);
        const double* h = c->hred.get();
        const double inter = std::fabs(h[58]);
        const double eta_max = inter > 0.0 ? (omega * h[56] + h[57] / omega) / (2.0 * inter) : INFINITY;
        const double kk = std::max(k, 1) + 1.0;  // k counts from 1: the first factor is 1 - 2^-0.3
        c->eta_ad = std::min((1.0 - std::pow(kk, -0.3)) * eta_max, (1.0 + std::pow(kk, -0.6)) * eta);
        if (eta <= eta_max || tries >= 64) return tries;
        launch_elem(n, FnCopy<float>{c->x.get(), c->xs.get()}, s);
        launch_elem(m, FnCopy<float>{c->y.get(), c->ys.get()}, s);
    }
}

// Candidate T(z) of the last iteration of a round with base (x_base, y_base)
// and scale factors (dp, dd), as a total point in FP64 (xtot, ytot).
void total_point(pm_ctx* c, double dp, double dd) {
    hipStream_t s = c->stream;
    launch_elem(c->n, FnTotalX{c->xbase.get(), c->d_dc.get(), c->xt.get(), 1.0 / dp, c->d_lx.get(), c->d_ux.get(),
                               c->xtot.get()},
                s);
    if (c->m > 0)
        launch_elem(c->m, FnTotalY{c->ybase.get(), c->d_dr.get(), c->yt.get(), 1.0 / dd, c->d_lr.get(),
                                   c->d_ur.get(), c->ytot.get()},
                    s);
}

// z <- T(z), anchor and last restart point too (restarts inside a round).
void restart_at_candidate(pm_ctx* c) {
    hipStream_t s = c->stream;
    for (float* d : {c->x.get(), c->x0.get(), c->xl.get()}) launch_elem(c->n, FnCopy<float>{d, c->xt.get()}, s);
    for (float* d : {c->y.get(), c->y0.get(), c->yl.get()}) launch_elem(c->m, FnCopy<float>{d, c->yt.get()}, s);
}

// The main solve's iterates, round data and round base into hold32 /
// hold64 (to_hold) or back.
void hold_state(pm_ctx* c, bool to_hold) {
    hipStream_t s = c->stream;
    const size_t n = c->n, m = std::max(c->m, 1);
    float* const xs[6] = {c->x.get(), c->x0.get(), c->xl.get(), c->c32.get(), c->lx32.get(), c->ux32.get()};
    float* const ys[6] = {c->y.get(), c->y0.get(), c->yl.get(), c->lr32.get(), c->ur32.get(), c->g32.get()};
    for (int q = 0; q < 6; ++q) {
        float* hx = c->hold32.get() + q * n;
        float* hy = c->hold32.get() + 6 * n + q * m;
        if (to_hold) {
            launch_elem(c->n, FnCopy<float>{hx, xs[q]}, s);
            launch_elem(c->m, FnCopy<float>{hy, ys[q]}, s);
        } else {
            launch_elem(c->n, FnCopy<float>{xs[q], hx}, s);
            launch_elem(c->m, FnCopy<float>{ys[q], hy}, s);
        }
    }
    double* hxb = c->hold64.get();
    double* hyb = c->hold64.get() + n;
    if (to_hold) {
        launch_elem(c->n, FnCopy<double>{hxb, c->xbase.get()}, s);
        launch_elem(c->m, FnCopy<double>{hyb, c->ybase.get()}, s);
    } else {
        launch_elem(c->n, FnCopy<double>{c->xbase.get(), hxb}, s);
        launch_elem(c->m, FnCopy<double>{c->ybase.get(), hyb}, s);
    }
}

// One phase of a feasibility polish (Polish in control.hpp) on the round
// data already built: reflected Halpern iterations with the main primal
// weight until the phase's residual (primal: rel_p of x_tot, dual: rel_d
// of y_tot) meets its target, restarts on it, budget.  Returns whether it
// met the target; the phase's total point is left in xtot / ytot.
bool polish_phase(pm_ctx* c, Control& ctl, Polish& pol, int phase, double scale, bool use_g, int K) {
    pol.start(phase, ctl.total_it);
    for (;;) {
        const int chunk = ctl.chunk(K);
        if (chunk <= 0) return false;
        hal_chunk(c, pol.k, ctl.omega, use_g, chunk, K);
        ctl.total_it += chunk;
        total_point(c, scale, scale);
        const Kkt kk = eval_kkt(c, c->xtot.get(), c->ytot.get());
        const int pa = pol.on_check(ctl.tol, phase == Polish::PRIMAL ? kk.rel_p : kk.rel_d, chunk);
        if (pa == Polish::DONE) return true;
        if (pa == Polish::FAILED) return false;
        if (pa == Polish::RESTART) restart_at_candidate(c);
    }
}

// Feasibility polish from the candidate (xtot, ytot) of errors kk: the
// primal phase from x_cand, the dual phase from y_cand, then the pair.
// Returns true when the pair is accepted (it is left in xtot / ytot and
// *out holds its KKT); otherwise the main solve's state is restored as it
// was and only the iterations spent remain counted.
bool polish(pm_ctx* c, Control& ctl, Polish& pol, const Kkt& kk, int K, Kkt* out) {
    hipStream_t s = c->stream;
    const int m = c->m, n = c->n;
    const double tol = ctl.tol;
    const Control held = ctl;
    hold_state(c, true);
    launch_elem(n, FnCopy<double>{c->xsave.get(), c->xtot.get()}, s);
    if (m > 0) launch_elem(m, FnCopy<double>{c->ysave.get(), c->ytot.get()}, s);
    pol.save(ctl.total_it, kk.rel, kk.rel_p, kk.rel_d);
    bool ok = true;
    if (pol.needs(tol, kk.rel_p)) {
        const double f = Polish::scale(kk.rel_p, ctl.DP);
        launch_elem(n, FnCopy<double>{c->xbase.get(), c->xsave.get()}, s);
        launch_elem(std::max(m, 1), FnFill<double>{c->ybase.get(), 0.0}, s);
        round_data(c, f, f, kRoundPrimalFeas);
        ok = polish_phase(c, ctl, pol, Polish::PRIMAL, f, false, K);
        if (ok) launch_elem(n, FnCopy<double>{c->xpol.get(), c->xtot.get()}, s);
    } else {
        launch_elem(n, FnCopy<double>{c->xpol.get(), c->xsave.get()}, s);
    }
    if (ok && pol.needs(tol, kk.rel_d)) {
        const double f = Polish::scale(kk.rel_d, ctl.DP);
        launch_elem(n, FnFill<double>{c->xbase.get(), 0.0}, s);
        if (m > 0) launch_elem(m, FnCopy<double>{c->ybase.get(), c->ysave.get()}, s);
        round_data(c, f, f, kRoundDualFeas);
        ok = polish_phase(c, ctl, pol, Polish::DUAL, f, true, K);
    } else if (ok && m > 0) {
        launch_elem(m, FnCopy<double>{c->ytot.get(), c->ysave.get()}, s);
    }
    pol.phase = Polish::MAIN;
    if (ok) {
        launch_elem(n, FnCopy<double>{c->xtot.get(), c->xpol.get()}, s);
        const Kkt pk = eval_kkt(c, c->xtot.get(), c->ytot.get());
        if (Polish::accept(tol, pk.rel, pk.rel_g)) {
            *out = pk;
            return true;
        }
    }
    const int spent = ctl.total_it;
    ctl = held;
    ctl.total_it = spent;
    ++pol.fails;
    hold_state(c, false);
    launch_elem(n, FnCopy<double>{c->xtot.get(), c->xsave.get()}, s);
    if (m > 0) launch_elem(m, FnCopy<double>{c->ytot.get(), c->ysave.get()}, s);
    return false;
}

// The normalized duality gap of the candidate (xt, yt) of the current
// round on the ball around it of radius ||(xt, yt) - (x0, y0)||_omega
// (FnNdgX, FnNdgY): three passes of kNdgT step scales each (a logarithmic
// sweep, then two linear refinements of the bracket where the distance
// crosses the radius), the gap interpolated at the crossing.  Returns -1
// (use the KKT error) when the radius is 0.
double normalized_gap(pm_ctx* c, double omega, bool use_g) {
    hipStream_t s = c->stream;
    const int m = c->m, n = c->n;
    double* red = c->red.get();
    launch_elem_reduce(n, FnNormDiff{c->xt.get(), c->x0.get()}, c->partials.get(), red + 50, s);
    launch_elem_reduce(m, FnNormDiff{c->yt.get(), c->y0.get()}, c->partials.get(), red + 51, s);
    launch_rows<float>(c->var_at, n, c->cp.get(), c->ri.get(), c->vat.get(), c->yt.get(),
                       EpiStore<float>{c->tmp_n.get()}, s);
    if (m > 0)
        launch_rows<float>(c->var_a, m, c->rp.get(), c->ci.get(), c->va.get(), c->xt.get(),
                           EpiStore<float>{c->tmp_m.get()}, s);
    launch_elem_reduce(n, FnNdgScale{c->tmp_n.get(), c->c32.get()}, c->partials.get(), red + 52, s);
    launch_elem_reduce(m, FnNdgScale{c->tmp_m.get(), nullptr}, c->partials.get(), red + 53, s);
    PM_HIP(hipMemcpyAsync(c->hred.get() + 50, red + 50, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    sync(c);
    const double* h = c->hred.get();
    const double r = std::sqrt(omega * h[50] + h[51] / omega);
    if (!(r > 0.0) || !std::isfinite(r)) return -1.0;
    const double t0 = r / std::sqrt(h[52] / omega + omega * h[53] + 1e-300);
    double nrm[kNdgT], val[kNdgT];
    auto pass = [&](const NdgTs& ts) {
        double* out = c->ndg_red.get();
        launch_elem_reduce(n, FnNdgX{c->xt.get(), c->tmp_n.get(), c->c32.get(), c->lx32.get(), c->ux32.get(), omega, ts},
                           c->ndg_part.get(), out, s);
        launch_elem_reduce(m, FnNdgY{c->yt.get(), c->tmp_m.get(), c->lr32.get(), c->ur32.get(),
                                     use_g ? c->g32.get() : nullptr, omega, ts},
                           c->ndg_part.get(), out + 2 * kNdgT, s);
        PM_HIP(hipMemcpyAsync(c->ndg_h.get(), out, 4 * kNdgT * sizeof(double), hipMemcpyDeviceToHost, s));
        sync(c);
        const double* o = c->ndg_h.get();
        for (int k = 0; k < kNdgT; ++k) {
            nrm[k] = std::sqrt(o[2 * k] + o[2 * kNdgT + 2 * k]);
            val[k] = o[2 * k + 1] + o[2 * kNdgT + 2 * k + 1];
        }
    };
    NdgTs ts;
    for (int k = 0; k < kNdgT; ++k) ts.t[k] = t0 * std::pow(10.0, k - 5);
    pass(ts);
    double lo_t = 0.0, lo_n = 0.0, lo_v = 0.0, hi_t = 0.0, hi_n = 0.0, hi_v = 0.0;
    for (int round = 0; round < 3; ++round) {
        int k = 0;
        while (k < kNdgT && nrm[k] < r) ++k;
        if (k == kNdgT) {
            if (round == 0) return val[kNdgT - 1] / r;  // the box ends before the ball: its far point
            lo_t = ts.t[kNdgT - 1]; lo_n = nrm[kNdgT - 1]; lo_v = val[kNdgT - 1];  // (hi unchanged)
        } else {
            if (k > 0) { lo_t = ts.t[k - 1]; lo_n = nrm[k - 1]; lo_v = val[k - 1]; }
            hi_t = ts.t[k]; hi_n = nrm[k]; hi_v = val[k];
        }
        if (round == 2) break;
        for (int q = 0; q < kNdgT; ++q) ts.t[q] = lo_t + (hi_t - lo_t) * (q + 1) / (kNdgT + 1);
        pass(ts);
    }
    const double w = hi_n > lo_n ? (r - lo_n) / (hi_n - lo_n) : 1.0;
    return (lo_v + w * (hi_v - lo_v)) / r;
}

void pdlp_solve(pm_ctx* c, pm_pdhg_result* res) {
    hipStream_t s = c->stream;
    const int m = c->m, n = c->n;
    const auto& o = c->opts;
    const int K = std::max(1, o.check_frequency);

    // Round 0 base point: the warm start (clamped to the bounds), or zero.
    if (!c->x0_h.empty()) {
        std::vector<double> xb(n);
        for (int j = 0; j < n; ++j) xb[j] = std::min(std::max(c->x0_h[j], c->lx[j]), c->ux[j]);
        c->xbase.upload(xb, s);
    } else {
        launch_elem(n, FnFill<double>{c->xbase.get(), 0.0}, s);
    }
    if (!c->y0_h.empty() && m > 0) c->ybase.upload(c->y0_h, s);
    else launch_elem(std::max(m, 1), FnFill<double>{c->ybase.get(), 0.0}, s);

    // A warm start is judged by its FP64 relative KKT error against the
    // cold start's (x = 0 clipped to the bounds, y = 0): the worse of the
    // two is not used.
    const bool warm_given = !c->x0_h.empty() || (!c->y0_h.empty() && m > 0);
    bool warm = warm_given;
    if (warm_given) {
        launch_elem(n, FnCopy<double>{c->xtot.get(), c->xbase.get()}, s);
        if (m > 0) launch_elem(m, FnCopy<double>{c->ytot.get(), c->ybase.get()}, s);
        const double rw = eval_kkt(c, c->xtot.get(), c->ytot.get()).rel;
        std::vector<double> xc(n);
        for (int j = 0; j < n; ++j) xc[j] = c->lx[j] > 0.0 ? c->lx[j] : (c->ux[j] < 0.0 ? c->ux[j] : 0.0);
        c->xtot.upload(xc, s);
        if (m > 0) launch_elem(m, FnFill<double>{c->ytot.get(), 0.0}, s);
        const double r0 = eval_kkt(c, c->xtot.get(), c->ytot.get()).rel;
        warm = rw <= r0;
        c->warm_accepted = warm ? 1 : 0;
        if (!warm) {
            launch_elem(n, FnFill<double>{c->xbase.get(), 0.0}, s);
            launch_elem(std::max(m, 1), FnFill<double>{c->ybase.get(), 0.0}, s);
        }
    } else {
        c->warm_accepted = -1;
    }

    Control ctl;
    ctl.init(o.tolerance, o.max_iterations, o.max_refinements, c->omega0, warm && !c->y0_h.empty() && m > 0);
    c->eta_ad = c->eta;  // the adaptive rule starts from the constant step
    res->status = PM_SOLVE_ITERATION_LIMIT;
    // Under PM_TERM_REFERENCE a point is optimal only when the reference's
    // own measures agree (pv, dv <= tolerance, pdhg_solver.cu:356): each
    // check that meets the relative target but not them tightens the
    // target tenfold and the solve goes on.
    const bool ref_test = o.termination == PM_TERM_REFERENCE;
    // Feasibility polishing stays off under the reference's test: a polished
    // primal point is feasible but not complementary entry by entry (values
    // slightly above a bound where the reduced cost is large), which the
    // reference's dv counts in full, so its pairs would be rejected.
    Polish pol;
    pol.init(o.feasibility_polishing != 0 && !ref_test);
    const auto t0 = std::chrono::steady_clock::now();
    bool have_candidate = false;
    bool round_start = true;
    // A solve that ends without meeting the tolerance returns the checked
    // candidate of lowest relative KKT error, not the last one: late in a
    // stalled solve the error can rise again after a restart.
    double best_rel = INFINITY;
    for (;;) {
        if (round_start) {
            round_data(c, ctl.DP, ctl.DD);
            round_start = false;
        }
        const int chunk = ctl.chunk(K);
        if (chunk <= 0) break;
        int extra = 0;  // rejected adaptive steps, each an iteration's work
        if (c->adaptive_step) {
            for (int i = 0; i < chunk; ++i)
                extra += adaptive_iteration(c, ctl.omega, i == chunk - 1, ctl.use_g, ctl.total_it + extra + i) - 1;
        } else {
            hal_chunk(c, ctl.k, ctl.omega, ctl.use_g, chunk, K);
        }
        ctl.advance(chunk);
        ctl.total_it += extra;
        // Candidate = T(z) of the last iteration, as a total point in FP64.
        total_point(c, ctl.DP, ctl.DD);
        have_candidate = true;
        const Kkt kk = eval_kkt(c, c->xtot.get(), c->ytot.get());
        if (kk.rel < best_rel) {
            best_rel = kk.rel;
            launch_elem(n, FnCopy<double>{c->xbest.get(), c->xtot.get()}, s);
            if (m > 0) launch_elem(m, FnCopy<double>{c->ybest.get(), c->ytot.get()}, s);
        }
        const double q = c->restart_on_gap && kk.rel > ctl.tol ? normalized_gap(c, ctl.omega, ctl.use_g) : -1.0;
        const int act = ctl.on_check(kk.rel, kk.rel_p, kk.rel_d, kk.rel_g, q);
        if (o.verbose) {
            fprintf(stderr, "[pramaan] it %d round %d kkt %.3e (p %.2e d %.2e g %.2e) omega %.3g restarts %d",
                    ctl.total_it, ctl.rounds, kk.rel, kk.rel_p, kk.rel_d, kk.rel_g, ctl.omega, ctl.restarts);
            if (q >= 0.0) fprintf(stderr, " ndg %.3e", q);
            if (c->adaptive_step) fprintf(stderr, " step %.3e (constant %.3e)", c->eta_ad, c->eta);
            fprintf(stderr, "\n");
        }
        if (act == Control::OPTIMAL) {
            double pv = 0.0, dv = 0.0;
            if (ref_test) ref_style_residuals(c, c->xtot.get(), c->ytot.get(), o.tolerance, pv, dv);
            if (!ref_test || (pv <= o.tolerance && dv <= o.tolerance)) {
                res->status = PM_SOLVE_OPTIMAL;
                break;
            }
            ctl.tol = std::max(ctl.tol * 0.1, 1e-14);
            continue;
        }
        if (act == Control::CONTINUE && pol.trigger(ctl.tol, ctl.total_it, kk.rel, kk.rel_g)) {
            Kkt pk;
            if (polish(c, ctl, pol, kk, K, &pk)) {
                if (o.verbose)
                    fprintf(stderr, "[pramaan] it %d polished kkt %.3e (p %.2e d %.2e g %.2e)\n", ctl.total_it, pk.rel,
                            pk.rel_p, pk.rel_d, pk.rel_g);
                res->status = PM_SOLVE_OPTIMAL;
                break;
            }
            continue;  // the main solve goes on where it was
        }
        if (act == Control::FAIL) { res->status = PM_SOLVE_NUMERICAL_FAILURE; break; }
        if (act == Control::LIMIT) break;
        if (act == Control::REFINE) {
            std::swap(c->xbase, c->xtot);
            std::swap(c->ybase, c->ytot);
            round_start = true;
            have_candidate = false;
            continue;
        }
        if (act == Control::RESTART) {
            // z <- T(z); anchor z0 <- z; primal weight from the distances moved.
            double* red = c->red.get();
            launch_elem_reduce(n, FnNormDiff{c->xt.get(), c->xl.get()}, c->partials.get(), red + 48, s);
            launch_elem_reduce(m, FnNormDiff{c->yt.get(), c->yl.get()}, c->partials.get(), red + 49, s);
            PM_HIP(hipMemcpyAsync(c->hred.get() + 48, red + 48, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            launch_elem(n, FnCopy<float>{c->x.get(), c->xt.get()}, s);
            launch_elem(n, FnCopy<float>{c->x0.get(), c->xt.get()}, s);
            launch_elem(n, FnCopy<float>{c->xl.get(), c->xt.get()}, s);
            launch_elem(m, FnCopy<float>{c->y.get(), c->yt.get()}, s);
            launch_elem(m, FnCopy<float>{c->y0.get(), c->yt.get()}, s);
            launch_elem(m, FnCopy<float>{c->yl.get(), c->yt.get()}, s);
            sync(c);
            ctl.on_restart(std::sqrt(c->hred[48]), std::sqrt(c->hred[49]));
        }
    }
    sync(c);
    res->device_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->iterations = ctl.total_it;
    res->restarts = ctl.restarts;
    res->refinements = ctl.rounds;
    res->polish_attempts = pol.attempts;
    res->primal_weight = ctl.omega;
    if (res->status != PM_SOLVE_OPTIMAL && best_rel < INFINITY) {
        launch_elem(n, FnCopy<double>{c->xtot.get(), c->xbest.get()}, s);
        if (m > 0) launch_elem(m, FnCopy<double>{c->ytot.get(), c->ybest.get()}, s);
    } else if (!have_candidate) {
        // No iteration since the last base point: report the base point.
        launch_elem(n, FnCopy<double>{c->xtot.get(), c->xbase.get()}, s);
        if (m > 0) launch_elem(m, FnCopy<double>{c->ytot.get(), c->ybase.get()}, s);
    }
    std::vector<double> xh(n), yh(m);
    c->xtot.download(xh.data(), n, s);
    if (m > 0) c->ytot.download(yh.data(), m, s);
    sync(c);
    double pv = 0, dv = 0;
    ref_style_residuals(c, c->xtot.get(), c->ytot.get(), o.tolerance, pv, dv);
    res->primal_residual = pv;
    res->dual_residual = dv;
    finish_result(c, res, c->xtot.get(), c->ytot.get(), xh, yh);
}

void ctx_solve(pm_ctx* c, pm_pdhg_result* res) {
    if (!res) fail(PM_ERR_INVALID_ARG, "result is NULL");
    if (!c->loaded) fail(PM_ERR_INVALID_ARG, "no LP loaded");
    res->status = PM_SOLVE_ITERATION_LIMIT;
    res->iterations = 0;
    res->restarts = 0;
    res->refinements = 0;
    res->polish_attempts = 0;
    res->objective = 0.0;
    res->primal_residual = INFINITY;
    res->dual_residual = INFINITY;
    res->rel_kkt = res->rel_primal = res->rel_dual = res->rel_gap = INFINITY;
    res->primal_objective = res->dual_objective = 0.0;
    res->primal_weight = 0.0;
    res->device_seconds = 0.0;
    const int m = c->m, n = c->n;
    if (n == 0) {  // pdhg_solver.cu:178-184
        res->status = PM_SOLVE_OPTIMAL;
        res->objective = c->offset;
        if (res->row_activity) std::fill(res->row_activity, res->row_activity + m, 0.0);
        if (res->y) std::fill(res->y, res->y + m, 0.0);
        return;
    }
    if (c->nan_cost) {  // pdhg_solver.cu:186-192
        res->status = PM_SOLVE_NUMERICAL_FAILURE;
        return;
    }
    if (c->opts.algorithm == PM_ALG_REFERENCE) ref_solve(c, res);
    else pdlp_solve(c, res);
}

}  // namespace
}  // namespace pm

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* pm_version(void) { return "pramaan-mi355x 0.1 (gfx950)"; }

const char* pm_last_error(void) { return pm::g_last_error.c_str(); }

int pm_device_count(int32_t* count) {
    if (!count) return PM_ERR_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    (void)hipGetLastError();
    *count = c;
    return PM_OK;
}

void pm_pdhg_options_default(pm_pdhg_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tolerance = 1e-6;
    o->max_iterations = 100000;
    o->check_frequency = 100;
    o->algorithm = PM_ALG_REFERENCE;
    o->termination = PM_TERM_REFERENCE;
    o->spmv_variant = PM_SPMV_AUTO;
    o->ruiz_iterations = 10;
    o->max_refinements = 8;
    o->device = 0;
    o->verbose = 0;
    o->primal_weight = 0.0;
    o->x0 = nullptr;
    o->y0 = nullptr;
    o->feasibility_polishing = 1;
}

int pm_ctx_create(int32_t device, pm_ctx** out) {
    return pm::guarded([&] {
        if (!out) pm::fail(PM_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            (void)hipGetLastError();
            pm::fail(PM_ERR_NO_DEVICE, "no HIP device visible");
        }
        if (device < 0 || device >= count) pm::fail(PM_ERR_NO_DEVICE, "device ordinal out of range");
        PM_HIP(hipSetDevice(device));
        auto c = std::make_unique<pm_ctx>();
        c->device = device;
        PM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        PM_HIP(hipEventCreate(&c->ev0));
        PM_HIP(hipEventCreate(&c->ev1));
        *out = c.release();
    });
}

int pm_ctx_destroy(pm_ctx* c) {
    return pm::guarded([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        pm::drop_chunk_graph(c);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        hipStream_t s = c->stream;
        delete c;
        if (s) (void)hipStreamDestroy(s);
    });
}

int pm_ctx_set_graph(pm_ctx* c, int32_t use_graph) {
    return pm::guarded([&] {
        if (!c) pm::fail(PM_ERR_INVALID_ARG, "ctx is NULL");
        c->use_graph = use_graph != 0;
    });
}

int pm_ctx_load(pm_ctx* c, const pm_lp* lp, const pm_pdhg_options* opts) {
    pm::Range range_("pm_ctx_load");
    return pm::guarded([&] {
        if (!c) pm::fail(PM_ERR_INVALID_ARG, "ctx is NULL");
        PM_HIP(hipSetDevice(c->device));
        pm::ctx_load(c, lp, opts);
    });
}

int pm_ctx_solve(pm_ctx* c, pm_pdhg_result* res) {
    pm::Range range_("pm_ctx_solve");
    return pm::guarded([&] {
        if (!c) pm::fail(PM_ERR_INVALID_ARG, "ctx is NULL");
        PM_HIP(hipSetDevice(c->device));
        pm::ctx_solve(c, res);
    });
}

int pm_pdhg_solve(const pm_lp* lp, const pm_pdhg_options* opts, pm_pdhg_result* res) {
    pm::Range range_("pm_pdhg_solve");
    // The LP and the options are checked on the host before a device is
    // touched, so bad arguments fail the same way with or without a GPU.
    int rc = pm::guarded([&] {
        pm::check_lp(lp);
        pm_pdhg_options o;
        pm_pdhg_options_default(&o);
        if (opts) o = *opts;
        pm::check_options(o);
        if (!res) pm::fail(PM_ERR_INVALID_ARG, "result is NULL");
    });
    if (rc != PM_OK) return rc;
    pm_ctx* c = nullptr;
    const int dev = opts ? opts->device : 0;
    rc = pm_ctx_create(dev, &c);
    if (rc != PM_OK) return rc;
    rc = pm_ctx_load(c, lp, opts);
    if (rc == PM_OK) rc = pm_ctx_solve(c, res);
    const std::string err = pm::g_last_error;
    pm_ctx_destroy(c);
    if (rc != PM_OK) pm::set_last_error(err);
    return rc;
}

int pm_ctx_scaling(pm_ctx* c, double* row_scale, double* col_scale) {
    return pm::guarded([&] {
        if (!c || !row_scale || !col_scale) pm::fail(PM_ERR_INVALID_ARG, "NULL argument");
        if (!c->loaded || c->n == 0 || c->nan_cost) pm::fail(PM_ERR_INVALID_ARG, "no LP loaded");
        PM_HIP(hipSetDevice(c->device));
        if (c->m > 0) c->d_dr.download(row_scale, c->m, c->stream);
        c->d_dc.download(col_scale, c->n, c->stream);
        PM_HIP(hipStreamSynchronize(c->stream));
    });
}

int pm_ctx_spmv_variants(pm_ctx* c, int32_t* va, int32_t* vat) {
    return pm::guarded([&] {
        if (!c || !va || !vat) pm::fail(PM_ERR_INVALID_ARG, "NULL argument");
        *va = c->nblk > 0 ? PM_SPMV_COLBLOCK : c->merge_a ? PM_SPMV_MERGE : c->sell_a.slices > 0 ? PM_SPMV_SLICED : c->var_a;
        *vat = c->merge_at ? PM_SPMV_MERGE : c->sell_at.slices > 0 ? PM_SPMV_SLICED : c->var_at;
    });
}

int pm_ctx_time_kernel(pm_ctx* c, int32_t which, int32_t reps, double* avg_ms, double* alg_bytes) {
    return pm::guarded([&] {
        using namespace pm;
        if (!c || !avg_ms || !alg_bytes) fail(PM_ERR_INVALID_ARG, "NULL argument");
        if (!c->loaded || c->n == 0 || c->nan_cost) fail(PM_ERR_INVALID_ARG, "no LP loaded");
        if (reps <= 0) fail(PM_ERR_INVALID_ARG, "reps must be positive");
        PM_HIP(hipSetDevice(c->device));
        const double m = c->m, n = c->n, nnz = c->nnz;
        const double bytes_a = 8.0 * nnz + 4.0 * (m + 1) + 4.0 * n + 20.0 * m;
        const double bytes_at = 8.0 * nnz + 4.0 * (n + 1) + 4.0 * m + 28.0 * n;
        const bool pdlp = c->opts.algorithm == PM_ALG_PDLP;
        const double omega = c->omega0;
        auto launch = [&](int w) {
            if (w == PM_KERNEL_SPMV_A || w == PM_KERNEL_ITERATION) {
                if (pdlp) {
                    EpiDualHal ed{static_cast<float>(c->eta * omega), 0.5f, 1.0f, c->lr32.get(), c->ur32.get(),
                                  nullptr, c->y0.get(), c->y.get(), nullptr};
                    launch_dual_half(c, ed, c->stream);
                } else {
                    launch_rows<float>(c->var_a, c->m, c->rp.get(), c->ci.get(), c->va.get(), c->xe.get(),
                                       EpiDualRef{c->sig32.get(), c->lr32.get(), c->ur32.get(), c->y.get()}, c->stream);
                }
            }
            if (w == PM_KERNEL_SPMV_AT || w == PM_KERNEL_ITERATION) {
                if (pdlp) {
                    EpiPrimalHal ep{static_cast<float>(c->eta / omega), 0.5f, 1.0f, c->c32.get(), c->lx32.get(),
                                    c->ux32.get(), c->x0.get(), c->x.get(), c->xe.get(), nullptr};
                    launch_primal_half(c, ep, c->stream);
                } else {
                    launch_rows<float>(c->var_at, c->n, c->cp.get(), c->ri.get(), c->vat.get(), c->y.get(),
                                       EpiPrimalRef{c->tau32.get(), c->c32.get(), c->lx32.get(), c->ux32.get(),
                                                    c->x.get(), c->xe.get()},
                                       c->stream);
                }
            }
        };
        if (which < 0 || which > 2) fail(PM_ERR_INVALID_ARG, "unknown kernel id");
        // Inputs of the timed kernels must be finite: give the working
        // vectors a neutral state when no solve ran yet.
        launch(which);
        launch(which);
        PM_HIP(hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; ++r) launch(which);
        PM_HIP(hipEventRecord(c->ev1, c->stream));
        PM_HIP(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *avg_ms = ms / reps;
        *alg_bytes = which == PM_KERNEL_SPMV_A ? bytes_a : which == PM_KERNEL_SPMV_AT ? bytes_at : bytes_a + bytes_at;
    });
}

}  // extern "C"
