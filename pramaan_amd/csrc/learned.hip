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

// Per LP: scaled values, global indices and features.
template <class VT>
__global__ __launch_bounds__(kLB) void k_l_graph(const BatchDesc* __restrict__ desc, const int* __restrict__ rp,
                                                 const int* __restrict__ ci, const VT* __restrict__ oa,
                                                 const int* __restrict__ cp, const int* __restrict__ ri,
                                                 const VT* __restrict__ oat, const double* __restrict__ c,
                                                 const double* __restrict__ lr, const double* __restrict__ ur,
                                                 const double* __restrict__ lx, const double* __restrict__ ux,
                                                 const double* __restrict__ scal, int period_cols, int period_rows,
                                                 int* __restrict__ gci, float* __restrict__ ar,
                                                 int* __restrict__ gri, float* __restrict__ ac,
                                                 float* __restrict__ xr, float* __restrict__ xc) {
    const BatchDesc d = desc[blockIdx.x];
    const double s = scal[4 * blockIdx.x], cs = scal[4 * blockIdx.x + 1];
    const double maxc = scal[4 * blockIdx.x + 2], maxr = scal[4 * blockIdx.x + 3];
    const int m = d.m, n = d.n;
    const int pc = period_cols > 0 ? max(1, n / period_cols - 1) : 1;
    const int pr = period_rows > 0 ? max(1, m / period_rows - 1) : 1;
    for (int i = threadIdx.x; i < m; i += kLB) {
        const int b = rp[d.rp_off + i], e = rp[d.rp_off + i + 1];
        double mx = 0.0;
        for (int k = b; k < e; ++k) mx = fmax(mx, fabs(static_cast<double>(oa[d.nz_off + k])));
        const double den = mx > 0.0 ? mx : 1.0;
        for (int k = b; k < e; ++k) {
            ar[d.nz_off + k] = static_cast<float>(static_cast<double>(oa[d.nz_off + k]) / den);
            gci[d.nz_off + k] = static_cast<int>(d.col_off) + ci[d.nz_off + k];
        }
        const int64_t r = d.row_off + i;
        const double lo = lr[r], hi = ur[r];
        float* f = xr + r * kF;
        f[0] = static_cast<float>(fin(lo) ? lo / s : 0.0);
        f[1] = static_cast<float>(fin(hi) ? hi / s : 0.0);
        f[2] = lo > -1e29 ? 1.0f : 0.0f;
        f[3] = hi < 1e29 ? 1.0f : 0.0f;
        f[4] = lo == hi ? 1.0f : 0.0f;
        f[5] = static_cast<float>(static_cast<double>(e - b) / maxr);
        f[6] = period_rows > 0 ? static_cast<float>(static_cast<double>(i / period_rows) / pr) : 0.0f;
        f[7] = 1.0f;
    }
    for (int j = threadIdx.x; j < n; j += kLB) {
        const int b = cp[d.cp_off + j], e = cp[d.cp_off + j + 1];
        double mx = 0.0;
        for (int k = b; k < e; ++k) mx = fmax(mx, fabs(static_cast<double>(oat[d.nz_off + k])));
        const double den = mx > 0.0 ? mx : 1.0;
        for (int k = b; k < e; ++k) {
            ac[d.nz_off + k] = static_cast<float>(static_cast<double>(oat[d.nz_off + k]) / den);
            gri[d.nz_off + k] = static_cast<int>(d.row_off) + ri[d.nz_off + k];
        }
        const int64_t q = d.col_off + j;
        const double lo = lx[q], hi = ux[q];
        float* f = xc + q * kF;
        f[0] = static_cast<float>(c[q] / cs);
        f[1] = static_cast<float>(fin(lo) ? lo / s : 0.0);
        f[2] = static_cast<float>(fin(hi) ? hi / s : 0.0);
        f[3] = lo > -1e29 ? 1.0f : 0.0f;
        f[4] = hi < 1e29 ? 1.0f : 0.0f;
        f[5] = static_cast<float>(static_cast<double>(e - b) / maxc);
        f[6] = period_cols > 0 ? static_cast<float>(static_cast<double>(j / period_cols) / pc) : 0.0f;
        f[7] = 1.0f;
    }
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
        hipLaunchKernelGGL(k_l_graph<VT>, dim3(count), dim3(kLB), 0, s, b->d_desc.get(), b->rp.get(), b->ci.get(),
                           oa, b->cp.get(), b->ri.get(), oat, b->c_model_dev.get(), b->lr.get(), b->ur.get(),
                           b->lx.get(), b->ux.get(), scal.get(), period_cols, period_rows, g->ci.get(), g->ar.get(),
                           g->ri.get(), g->ac.get(), g->xr.get(), g->xc.get());
    };
    if (b->exact) graph(static_cast<const float*>(b->oa32.get()), static_cast<const float*>(b->oat32.get()));
    else graph(static_cast<const double*>(b->oa64.get()), static_cast<const double*>(b->oat64.get()));
    PM_LAUNCHED();
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
    for (int k = 0; k < count; ++k) {
        b->desc[k].warm_x = 1;
        b->desc[k].warm_y = b->desc[k].m > 0 ? 1 : 0;
    }
}

}  // namespace
}  // namespace pm

extern "C" {

int pm_batch_learned_warm_start(pm_batch* b, pm_gnn* g, int32_t period_cols, int32_t period_rows,
                                double* seconds) {
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
