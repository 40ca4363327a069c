// learned.hip -- learned warm start of a resident batch: the network's
// graph and features of every LP are built on the device from the batch's
// own arrays (no host round trip), one forward pass runs over the
// block-diagonal graph of all LPs, and the predictions become the batch's
// warm starts.  The features are those of pm_gnn_lp_features (host), value
// for value: the same double-precision expressions rounded once to float.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "batch_internal.hpp"
#include "gnn_internal.hpp"
#include "pramaan_gnn.h"
#include "pramaan_hip.h"

namespace pm {
namespace {

constexpr int kLB = 1024;  // one workgroup per LP
constexpr int kF = PM_GNN_FEATURES;

__device__ __forceinline__ bool fin(double v) { return fabs(v) < 1e29; }

// Block-wide maximum of four values per thread, result in out[0..3].
__device__ void block_max4(double (&v)[4], double* out) {
    __shared__ double tmp[kLB / 64][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double a = v[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(a, o, 64);
            a = t > a ? t : a;
        }
        v[j] = a;
    }
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 4; ++j) tmp[threadIdx.x >> 6][j] = v[j];
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = tmp[0][threadIdx.x];
        for (int w = 1; w < kLB / 64; ++w) a = tmp[w][threadIdx.x] > a ? tmp[w][threadIdx.x] : a;
        out[threadIdx.x] = a;
    }
    __syncthreads();
}

// Per LP: s, cs, largest column and row length -> scal[4 lp ..].
__global__ __launch_bounds__(kLB) void k_l_scales(const BatchDesc* __restrict__ desc, const int* __restrict__ rp,
                                                  const int* __restrict__ cp, const double* __restrict__ c,
                                                  const double* __restrict__ lr, const double* __restrict__ ur,
                                                  const double* __restrict__ lx, const double* __restrict__ ux,
                                                  double* __restrict__ scal) {
    const BatchDesc d = desc[blockIdx.x];
    const int tid = threadIdx.x;
    __shared__ double out[4];
    double v[4] = {0.0, 0.0, 1.0, 1.0};
    for (int i = tid; i < d.m; i += kLB) {
        const int64_t r = d.row_off + i;
        if (fin(lr[r])) v[0] = fmax(v[0], fabs(lr[r]));
        if (fin(ur[r])) v[0] = fmax(v[0], fabs(ur[r]));
        v[3] = fmax(v[3], static_cast<double>(rp[d.rp_off + i + 1] - rp[d.rp_off + i]));
    }
    for (int j = tid; j < d.n; j += kLB) {
        const int64_t q = d.col_off + j;
        if (fin(lx[q])) v[0] = fmax(v[0], fabs(lx[q]));
        if (fin(ux[q])) v[0] = fmax(v[0], fabs(ux[q]));
        v[1] = fmax(v[1], fabs(c[q]));
        v[2] = fmax(v[2], static_cast<double>(cp[d.cp_off + j + 1] - cp[d.cp_off + j]));
    }
    block_max4(v, out);
    if (tid == 0) {
        scal[4 * blockIdx.x + 0] = fmax(1.0, out[0]);
        scal[4 * blockIdx.x + 1] = d.n > 0 ? fmax(1e-12, out[1]) : 1.0;
        scal[4 * blockIdx.x + 2] = out[2];
        scal[4 * blockIdx.x + 3] = out[3];
    }
}

// LP that owns global row (COLS: column) q: the last LP whose first row
// (column) is at most q.  Offsets grow with the LP index; an LP without
// rows shares its offset with the next one and owns none of them.
template <bool COLS>
__device__ __forceinline__ int lp_of(const BatchDesc* __restrict__ desc, int count, int64_t q) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((COLS ? desc[mid].col_off : desc[mid].row_off) <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The graph and features of every row (COLS: column) of the batch, L lanes
// per row: the lanes take the row's largest |a| (a maximum, exact in any
// order), write its entries scaled by it with their global indexes, and the
// first lane writes the row's features.  Rows of A give the row nodes'
// edges (ar, gci) and features (xr), rows of A^T the column nodes'.
template <class VT, int L, bool COLS>
__global__ __launch_bounds__(256) void k_l_side(int total, int count, const BatchDesc* __restrict__ desc,
                                                const int* __restrict__ ptr, const int* __restrict__ idx,
                                                const VT* __restrict__ val, const double* __restrict__ c,
                                                const double* __restrict__ lo_b, const double* __restrict__ hi_b,
                                                const double* __restrict__ scal, int period,
                                                int* __restrict__ gidx, float* __restrict__ a,
                                                float* __restrict__ feat) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t q = t / L;
    const int lane = static_cast<int>(t % L);
    const bool on = q < total;
    int k = 0, b = 0, e = 0;
    int64_t nz = 0;
    double mx = 0.0;
    if (on) {
        k = lp_of<COLS>(desc, count, q);
        const int64_t local = q - (COLS ? desc[k].col_off : desc[k].row_off);
        const int64_t po = (COLS ? desc[k].cp_off : desc[k].rp_off) + local;
        nz = desc[k].nz_off;
        b = ptr[po];
        e = ptr[po + 1];
        for (int kk = b + lane; kk < e; kk += L) mx = fmax(mx, fabs(static_cast<double>(val[nz + kk])));
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, L));
    if (!on) return;
    const BatchDesc& d = desc[k];
    const double den = mx > 0.0 ? mx : 1.0;
    const int base = static_cast<int>(COLS ? d.row_off : d.col_off);
    for (int kk = b + lane; kk < e; kk += L) {
        a[nz + kk] = static_cast<float>(static_cast<double>(val[nz + kk]) / den);
        gidx[nz + kk] = base + idx[nz + kk];
    }
    if (lane != 0) return;
    const int i = static_cast<int>(q - (COLS ? d.col_off : d.row_off));
    const double s = scal[4 * k], lo = lo_b[q], hi = hi_b[q];
    float* f = feat + q * kF;
    if constexpr (COLS) {
        const double cs = scal[4 * k + 1], maxc = scal[4 * k + 2];
        const int pc = period > 0 ? max(1, d.n / period - 1) : 1;
        f[0] = static_cast<float>(c[q] / cs);
        f[1] = static_cast<float>(fin(lo) ? lo / s : 0.0);
        f[2] = static_cast<float>(fin(hi) ? hi / s : 0.0);
        f[3] = lo > -1e29 ? 1.0f : 0.0f;
        f[4] = hi < 1e29 ? 1.0f : 0.0f;
        f[5] = static_cast<float>(static_cast<double>(e - b) / maxc);
        f[6] = period > 0 ? static_cast<float>(static_cast<double>(i / period) / pc) : 0.0f;
    } else {
        const double maxr = scal[4 * k + 3];
        const int pr = period > 0 ? max(1, d.m / period - 1) : 1;
        f[0] = static_cast<float>(fin(lo) ? lo / s : 0.0);
        f[1] = static_cast<float>(fin(hi) ? hi / s : 0.0);
        f[2] = lo > -1e29 ? 1.0f : 0.0f;
        f[3] = hi < 1e29 ? 1.0f : 0.0f;
        f[4] = lo == hi ? 1.0f : 0.0f;
        f[5] = static_cast<float>(static_cast<double>(e - b) / maxr);
        f[6] = period > 0 ? static_cast<float>(static_cast<double>(i / period) / pr) : 0.0f;
    }
    const int role = period > 0 ? i % period : -1;  // bits of the node's index within its period
#pragma unroll
    for (int bit = 0; bit < PM_GNN_ROLE_BITS; ++bit)
        f[7 + bit] = role < 0 ? 0.0f : (((role >> bit) & 1) ? 1.0f : -1.0f);
}

// Lanes per row for a mean row length: about two entries per lane.
int side_lanes(double mean) {
    int l = 1;
    while (l < 32 && 2.0 * l < mean) l *= 2;
    return l;
}

template <class VT, bool COLS>
void launch_side(hipStream_t s, int lanes, int total, int count, const BatchDesc* desc, const int* ptr,
                 const int* idx, const VT* val, const double* c, const double* lo, const double* hi,
                 const double* scal, int period, int* gidx, float* a, float* feat) {
    if (total <= 0) return;
    auto go = [&](auto lc) {
        constexpr int L = decltype(lc)::value;
        const dim3 grid(static_cast<unsigned>((static_cast<int64_t>(total) * L + 255) / 256));
        hipLaunchKernelGGL((k_l_side<VT, L, COLS>), grid, dim3(256), 0, s, total, count, desc, ptr, idx, val, c,
                           lo, hi, scal, period, gidx, a, feat);
    };
    switch (lanes) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 16: go(std::integral_constant<int, 16>{}); break;
        default: go(std::integral_constant<int, 32>{}); break;
    }
    PM_LAUNCHED();
}

// Predictions -> warm starts: x = clamp(xhat * s, l, u), y = yhat * cs
// (the batch's init phase projects y on the sign each row allows).
__global__ __launch_bounds__(kLB) void k_l_apply(const BatchDesc* __restrict__ desc, const double* __restrict__ scal,
                                                 const float* __restrict__ xhat, const float* __restrict__ yhat,
                                                 const double* __restrict__ lx, const double* __restrict__ ux,
                                                 double* __restrict__ wx, double* __restrict__ wy) {
    const BatchDesc d = desc[blockIdx.x];
    const double s = scal[4 * blockIdx.x], cs = scal[4 * blockIdx.x + 1];
    for (int j = threadIdx.x; j < d.n; j += kLB) {
        const int64_t q = d.col_off + j;
        const double v = static_cast<double>(xhat[q]) * s;
        wx[q] = v < lx[q] ? lx[q] : (v > ux[q] ? ux[q] : v);
    }
    for (int i = threadIdx.x; i < d.m; i += kLB) {
        const int64_t r = d.row_off + i;
        wy[r] = static_cast<double>(yhat[r]) * cs;
    }
}

void learned_warm_start(pm_batch* b, pm_gnn* g, int period_cols, int period_rows, double* seconds) {
    hipStream_t s = b->stream;
    const int count = b->count;
    const int M = static_cast<int>(b->rows_total), N = static_cast<int>(b->cols_total);
    const int Z = static_cast<int>(b->nnz_total);
    if (b->c_model_dev.size() != b->c_model.size()) b->c_model_dev.upload(b->c_model, s);
    DevBuf<double> scal(static_cast<size_t>(4) * count);
    // The network's graph buffers, sized for the whole batch.
    g->rp.resize(static_cast<size_t>(M) + 1);
    g->cp.resize(static_cast<size_t>(N) + 1);
    for (auto* p : {&g->ci, &g->ri}) p->resize(std::max(Z, 1));
    for (auto* p : {&g->ar, &g->ac}) p->resize(std::max(Z, 1));
    g->xr.resize(static_cast<size_t>(std::max(M, 1)) * kF);
    g->xc.resize(static_cast<size_t>(std::max(N, 1)) * kF);
    gnn_reserve(g, M, N, Z, false);
    std::vector<int> seg(static_cast<size_t>(count) + 1);  // each LP's columns, for the primal-weight head
    for (int k = 0; k < count; ++k) seg[k] = static_cast<int>(b->desc[k].col_off);
    seg[count] = N;
    gnn_set_segments(g, count, seg.data());
    b->d_desc.upload(b->desc, s);
    PM_HIP(hipEventRecord(b->ev0, s));
    PM_HIP(hipMemcpyAsync(g->rp.get(), b->grp.get(), sizeof(int) * (static_cast<size_t>(M) + 1),
                          hipMemcpyDeviceToDevice, s));
    PM_HIP(hipMemcpyAsync(g->cp.get(), b->gcp.get(), sizeof(int) * (static_cast<size_t>(N) + 1),
                          hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_l_scales, dim3(count), dim3(kLB), 0, s, b->d_desc.get(), b->rp.get(), b->cp.get(),
                       b->c_model_dev.get(), b->lr.get(), b->ur.get(), b->lx.get(), b->ux.get(), scal.get());
    PM_LAUNCHED();
    auto graph = [&](auto oa, auto oat) {
        using VT = std::remove_cv_t<std::remove_pointer_t<decltype(oa)>>;
        const BatchDesc* dd = b->d_desc.get();
        launch_side<VT, false>(s, side_lanes(M > 0 ? static_cast<double>(Z) / M : 0.0), M, count, dd, b->rp.get(),
                               b->ci.get(), oa, nullptr, b->lr.get(), b->ur.get(), scal.get(), period_rows,
                               g->ci.get(), g->ar.get(), g->xr.get());
        launch_side<VT, true>(s, side_lanes(N > 0 ? static_cast<double>(Z) / N : 0.0), N, count, dd, b->cp.get(),
                              b->ri.get(), oat, b->c_model_dev.get(), b->lx.get(), b->ux.get(), scal.get(),
                              period_cols, g->ri.get(), g->ac.get(), g->xc.get());
    };
    if (b->exact) graph(static_cast<const float*>(b->oa32.get()), static_cast<const float*>(b->oat32.get()));
    else graph(static_cast<const double*>(b->oa64.get()), static_cast<const double*>(b->oat64.get()));
    // The network runs on its own stream: order it after the features and
    // the warm-start writes after it.
    hipEvent_t ready = nullptr, done = nullptr;
    PM_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    PM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    try {
        PM_HIP(hipEventRecord(ready, s));
        PM_HIP(hipStreamWaitEvent(g->s, ready, 0));
        gnn_forward(g);
        PM_HIP(hipEventRecord(done, g->s));
        PM_HIP(hipStreamWaitEvent(s, done, 0));
    } catch (...) {
        (void)hipEventDestroy(ready);
        (void)hipEventDestroy(done);
        throw;
    }
    hipLaunchKernelGGL(k_l_apply, dim3(count), dim3(kLB), 0, s, b->d_desc.get(), scal.get(), g->xhat.get(),
                       g->yhat.get(), b->lx.get(), b->ux.get(), b->wx.get(), b->wy.get());
    PM_LAUNCHED();
    PM_HIP(hipEventRecord(b->ev1, s));
    PM_HIP(hipEventSynchronize(b->ev1));
    (void)hipEventDestroy(ready);
    (void)hipEventDestroy(done);
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (seconds) *seconds = ms * 1e-3;
    // The primal-weight head predicts log(final weight / default weight);
    // the factor is kept within the engine's span around the default.
    std::vector<float> what(count);
    g->what.download(what.data(), count, g->s);
    PM_HIP(hipStreamSynchronize(g->s));
    const double lim = std::log(Control::kOmegaSpan);
    for (int k = 0; k < count; ++k) {
        b->desc[k].warm_x = 1;
        b->desc[k].warm_y = b->desc[k].m > 0 ? 1 : 0;
        const double w = std::isfinite(what[k]) ? std::max(-lim, std::min(lim, static_cast<double>(what[k]))) : 0.0;
        b->desc[k].omega_scale = g->omega_head ? std::exp(w) : 0.0;
    }
}

}  // namespace
}  // namespace pm

extern "C" {

int pm_batch_learned_warm_start(pm_batch* b, pm_gnn* g, int32_t period_cols, int32_t period_rows,
                                double* seconds) {
    pm::Range range_("pm_batch_learned_warm_start");
    return pm::guarded([&] {
        if (!b || !g) pm::fail(PM_ERR_INVALID_ARG, "NULL batch or network");
        if (b->device != g->device) pm::fail(PM_ERR_INVALID_ARG, "batch and network live on different devices");
        if (period_cols < 0 || period_rows < 0) pm::fail(PM_ERR_INVALID_ARG, "negative period length");
        PM_HIP(hipSetDevice(b->device));
        pm::learned_warm_start(b, g, period_cols, period_rows, seconds);
    });
}

int pm_batch_get_warm_start(pm_batch* b, int32_t i, double* x0, double* y0) {
    return pm::guarded([&] {
        if (!b || i < 0 || i >= b->count) pm::fail(PM_ERR_INVALID_ARG, "bad batch or index");
        PM_HIP(hipSetDevice(b->device));
        const pm::BatchDesc& d = b->desc[i];
        if (!d.warm_x && !d.warm_y) pm::fail(PM_ERR_INVALID_ARG, "LP has no warm start");
        if (x0 && d.warm_x)
            PM_HIP(hipMemcpyAsync(x0, b->wx.get() + d.col_off, sizeof(double) * d.n, hipMemcpyDeviceToHost, b->stream));
        if (y0 && d.warm_y)
            PM_HIP(hipMemcpyAsync(y0, b->wy.get() + d.row_off, sizeof(double) * d.m, hipMemcpyDeviceToHost, b->stream));
        PM_HIP(hipStreamSynchronize(b->stream));
    });
}

}  // extern "C"
