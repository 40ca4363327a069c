// gnn.hip -- forward, backward and Adam kernels of the warm-start network
// (include/pramaan_gnn.h).
//
// Activations are N x 64 bfloat16 row-major (one 128-byte row per node),
// rounded to nearest even once, where they are produced; gradients are
// FP32.  The 64x64 dense layers run on gfx950's MFMA
// (v_mfma_f32_32x32x16_bf16, K = 16 per instruction): one wavefront
// computes a 32-row x 64-column output tile as two 32x32 accumulators,
// loading eight bfloat16 activations (or FP32 gradients, rounded as they
// are loaded) per lane and step, and reading weights from a pre-converted
// bfloat16 copy laid out k-contiguous per output column.  Weight gradients (64x64, summed over
// all nodes) use the same instruction with K running over nodes, one
// partial per workgroup, reduced in a fixed order.  Message passing is a
// sparse-times-dense product with eight lanes per output node and eight
// features per lane.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "gnn_internal.hpp"
#include "pramaan_gnn.h"

namespace pm {
namespace {

constexpr int kH = 64, kF = PM_GNN_FEATURES;
constexpr float kOmegaLossWeight = 0.01f;  // weight of the primal-weight head's term in the loss
static_assert(kF % 4 == 0, "features are loaded four at a time");
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<short>(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(static_cast<unsigned>(static_cast<unsigned short>(h)) << 16); }

// Eight consecutive values of a row as bfloat16 bits (16-byte load of an
// activation row, or two 16-byte loads of an FP32 gradient row rounded).
__device__ __forceinline__ s16x8 load8_bf(const short* p) { return *reinterpret_cast<const s16x8*>(p); }
__device__ __forceinline__ s16x8 load8_bf(const float* p) {
    const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
    s16x8 r;
    r[0] = f2bf(a.x); r[1] = f2bf(a.y); r[2] = f2bf(a.z); r[3] = f2bf(a.w);
    r[4] = f2bf(b.x); r[5] = f2bf(b.y); r[6] = f2bf(b.z); r[7] = f2bf(b.w);
    return r;
}
// Eight consecutive values of a row as FP32.
__device__ __forceinline__ void load8_f(const short* p, float (&v)[8]) {
    const s16x8 h = *reinterpret_cast<const s16x8*>(p);
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = bf2f(h[q]);
}
__device__ __forceinline__ void load8_f(const float* p, float (&v)[8]) {
    const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void store_val(short* p, float v) { *p = f2bf(v); }
__device__ __forceinline__ void store_val(float* p, float v) { *p = v; }
__device__ __forceinline__ float load_val(const short* p) { return bf2f(*p); }
__device__ __forceinline__ float load_val(const float* p) { return *p; }

// Row of a 32x32 MFMA accumulator element e held by `lane`
// (column = lane % 32).
__device__ __forceinline__ int acc_row(int e, int lane) { return 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3); }

// One 32x32x16 step: lane l supplies A[l % 32][8 (l / 32) .. + 7] and
// B[8 (l / 32) .. + 7][l % 32], eight bfloat16 values each.
__device__ __forceinline__ f32x16 mfma16(s16x8 a, s16x8 b, f32x16 acc) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
}

// Out[N][64] = act(A1 B1 (+ A2 B2) (+ bias)) (+ Out when ACC), with
// B given as BT[c][k] = B[k][c] in bfloat16.  Block = 4 waves = 128 rows.
// AT / OT: short (bfloat16 activations) or float (FP32 gradients).
template <bool TWO, bool BIAS, bool RELU, bool ACC, class AT, class OT>
__global__ __launch_bounds__(256) void k_dense(int N, const AT* __restrict__ A1, const short* __restrict__ B1T,
                                               const AT* __restrict__ A2, const short* __restrict__ B2T,
                                               const float* __restrict__ bias, OT* __restrict__ Out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + w) * 32;
    if (row0 >= N) return;
    const int r = row0 + (lane & 31), kg = lane >> 5, c = lane & 31;
    f32x16 acc0, acc1;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc0[e] = acc1[e] = 0.0f;
#pragma unroll
    for (int in = 0; in < (TWO ? 2 : 1); ++in) {
        const AT* A = in == 0 ? A1 : A2;
        const short* BT = in == 0 ? B1T : B2T;
#pragma unroll
        for (int ks = 0; ks < kH / 16; ++ks) {
            const int k0 = ks * 16 + 8 * kg;
            s16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
            if (r < N) a = load8_bf(A + static_cast<size_t>(r) * kH + k0);
            const s16x8 b0 = *reinterpret_cast<const s16x8*>(BT + c * kH + k0);
            const s16x8 b1 = *reinterpret_cast<const s16x8*>(BT + (32 + c) * kH + k0);
            acc0 = mfma16(a, b0, acc0);
            acc1 = mfma16(a, b1, acc1);
        }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int row = row0 + acc_row(e, lane);
        if (row >= N) continue;
        float v0 = acc0[e], v1 = acc1[e];
        if (BIAS) { v0 += bias[c]; v1 += bias[32 + c]; }
        if (RELU) { v0 = fmaxf(v0, 0.0f); v1 = fmaxf(v1, 0.0f); }
        OT* o = Out + static_cast<size_t>(row) * kH;
        if (ACC) { v0 += load_val(o + c); v1 += load_val(o + 32 + c); }
        store_val(o + c, v0);
        store_val(o + 32 + c, v1);
    }
}

// Partial weight gradient over rows [b*rows_per, (b+1)*rows_per):
// P[b][i][j] = sum_r A[r][i] D[r][j] (bfloat16 inputs, FP32 accumulate).
// Wave w computes the 32x32 tile (w / 2, w % 2).
__global__ __launch_bounds__(256) void k_wgrad(int N, int rows_per, const short* __restrict__ A,
                                               const float* __restrict__ D, float* __restrict__ P) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ti = w >> 1, tj = w & 1, kg = lane >> 5, c = lane & 31;
    const int rb = blockIdx.x * rows_per;
    int re = rb + rows_per;
    if (re > N) re = N;
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    for (int r0 = rb; r0 < re; r0 += 16) {
        s16x8 a, b;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = r0 + 8 * kg + q;
            const bool ok = r < re;
            a[q] = ok ? A[static_cast<size_t>(r) * kH + 32 * ti + c] : static_cast<short>(0);
            b[q] = f2bf(ok ? D[static_cast<size_t>(r) * kH + 32 * tj + c] : 0.0f);
        }
        acc = mfma16(a, b, acc);
    }
    float* out = P + static_cast<size_t>(blockIdx.x) * kH * kH;
#pragma unroll
    for (int e = 0; e < 16; ++e) out[(32 * ti + acc_row(e, lane)) * kH + 32 * tj + c] = acc[e];
}

// Partial column sums over row chunks: P[b][c] = sum_r D[r][c] (width 64),
// and optionally of X^T D for the kF-feature input layer: P2[b][k][c].
__global__ __launch_bounds__(256) void k_colsum(int N, int rows_per, const float* __restrict__ D,
                                                const float* __restrict__ X, float* __restrict__ P,
                                                float* __restrict__ P2) {
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;  // 4 row groups
    const int rb = blockIdx.x * rows_per;
    int re = rb + rows_per;
    if (re > N) re = N;
    float s = 0.0f, sx[kF];
#pragma unroll
    for (int k = 0; k < kF; ++k) sx[k] = 0.0f;
    for (int r = rb + g; r < re; r += 4) {
        const float d = D[static_cast<size_t>(r) * kH + c];
        s += d;
        if (X) {
#pragma unroll
            for (int k = 0; k < kF; ++k) sx[k] += X[static_cast<size_t>(r) * kF + k] * d;
        }
    }
    __shared__ float red[4][kH * (kF + 1)];
    red[g][c] = s;
#pragma unroll
    for (int k = 0; k < kF; ++k) red[g][kH + k * kH + c] = sx[k];
    __syncthreads();
    if (g == 0) {
        float t = red[0][c] + red[1][c] + red[2][c] + red[3][c];
        P[static_cast<size_t>(blockIdx.x) * kH + c] = t;
        if (X)
            for (int k = 0; k < kF; ++k) {
                const int o = kH + k * kH + c;
                P2[(static_cast<size_t>(blockIdx.x) * kF + k) * kH + c] = red[0][o] + red[1][o] + red[2][o] + red[3][o];
            }
    }
}

// out[e] = sum_b P[b][e] in a fixed order: a workgroup takes 16 outputs,
// and slice q (of 16) sums b = q, q + 16, ...; the 16 slice sums are then
// added in slice order.  Grid: kSumBlock(width) workgroups.
constexpr int kSumLanes = 16;
constexpr int kSumSlices = 16;
__global__ __launch_bounds__(256) void k_sum_partials(int nb, int width, const float* __restrict__ P,
                                                      float* __restrict__ out) {
    const int el = threadIdx.x % kSumLanes, q = threadIdx.x / kSumLanes;
    const int e = blockIdx.x * kSumLanes + el;
    float s = 0.0f;
    if (e < width)
        for (int b = q; b < nb; b += kSumSlices) s += P[static_cast<size_t>(b) * width + e];
    __shared__ float red[kSumSlices][kSumLanes];
    red[q][el] = s;
    __syncthreads();
    if (q == 0 && e < width) {
        float t = 0.0f;
#pragma unroll
        for (int r = 0; r < kSumSlices; ++r) t += red[r][el];
        out[e] = t;
    }
}

// Out[i][:] (+)= sum_k val[k] H[idx[k]][:], eight lanes per row (eight
// features each: one 16-byte load of a bfloat16 activation row, two of an
// FP32 gradient row), so a wavefront walks eight rows at once.  The grid
// is 8 x per blocks of 32 rows; block b runs on XCD b % 8, which takes the
// (b % 8)-th eighth of the rows in order, so the rows of a window (and the
// H rows they gather) stay in one XCD's L2.  Entries are taken eight at a
// time (indexes and values, then the gathers, then the FMAs in entry
// order), so each feature's sum is the same as one entry at a time and up
// to 64 gathers are in flight per wavefront.  IT / OT: short (bfloat16
// activations) or float (FP32 gradients).
template <bool ACC, class IT, class OT>
__global__ __launch_bounds__(256) void k_spmm(int nrows, int per, const int* __restrict__ rp,
                                              const int* __restrict__ idx, const float* __restrict__ val,
                                              const IT* __restrict__ Hin, OT* __restrict__ Out) {
    const int blk = (static_cast<int>(blockIdx.x) & 7) * per + (static_cast<int>(blockIdx.x) >> 3);
    const int row = blk * 32 + (threadIdx.x >> 3), f = (threadIdx.x & 7) * 8;
    if (row >= nrows) return;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int k1 = rp[row + 1];
    int k = rp[row];
    for (; k + 8 <= k1; k += 8) {
        int i[8];
        float v[8];
        float h[8][8];
#pragma unroll
        for (int u = 0; u < 8; ++u) { i[u] = idx[k + u]; v[u] = val[k + u]; }
#pragma unroll
        for (int u = 0; u < 8; ++u) load8_f(Hin + static_cast<size_t>(i[u]) * kH + f, h[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int q = 0; q < 8; ++q) s[q] = __builtin_fmaf(v[u], h[u][q], s[q]);
    }
    for (; k < k1; ++k) {
        float h[8];
        load8_f(Hin + static_cast<size_t>(idx[k]) * kH + f, h);
        const float vk = val[k];
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = __builtin_fmaf(vk, h[q], s[q]);
    }
    OT* o = Out + static_cast<size_t>(row) * kH + f;
    if constexpr (ACC) {
        float t[8];
        load8_f(o, t);
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] += t[q];
    }
    if constexpr (sizeof(OT) == 2) {
        s16x8 r;
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] = f2bf(s[q]);
        *reinterpret_cast<s16x8*>(o) = r;
    } else {
        reinterpret_cast<float4*>(o)[0] = make_float4(s[0], s[1], s[2], s[3]);
        reinterpret_cast<float4*>(o)[1] = make_float4(s[4], s[5], s[6], s[7]);
    }
}

// Input layer: Out = relu(X W + b), X [N][kF], W [kF][64] FP32, Out in
// bfloat16.  Sixteen lanes per row, four outputs each (one 8-byte store),
// so a wavefront writes four rows.
__global__ __launch_bounds__(256) void k_input(int N, const float* __restrict__ X, const float* __restrict__ W,
                                               const float* __restrict__ b, short* __restrict__ Out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t r = t >> 4;
    const int c = static_cast<int>(t & 15) * 4;
    if (r >= N) return;
    float xs[kF];
#pragma unroll
    for (int q = 0; q < kF / 4; ++q) {
        const float4 x4 = *reinterpret_cast<const float4*>(X + r * kF + 4 * q);
        xs[4 * q] = x4.x;
        xs[4 * q + 1] = x4.y;
        xs[4 * q + 2] = x4.z;
        xs[4 * q + 3] = x4.w;
    }
    const float4 b4 = *reinterpret_cast<const float4*>(b + c);
    float s[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
    for (int k = 0; k < kF; ++k) {
        const float4 w = *reinterpret_cast<const float4*>(W + k * kH + c);
        s[0] = __builtin_fmaf(xs[k], w.x, s[0]);
        s[1] = __builtin_fmaf(xs[k], w.y, s[1]);
        s[2] = __builtin_fmaf(xs[k], w.z, s[2]);
        s[3] = __builtin_fmaf(xs[k], w.w, s[3]);
    }
    typedef short s16x4 __attribute__((ext_vector_type(4)));
    s16x4 o;
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = f2bf(fmaxf(s[q], 0.0f));
    *reinterpret_cast<s16x4*>(Out + r * kH + c) = o;
}

// Primal-weight head, one workgroup per segment (LP) k: pool[k][c] = mean
// of H[j][c] over the segment's columns, what[k] = pool[k] . w + b.  Lane
// t takes features 8 (t % 8) .. 8 (t % 8) + 7 of rows j0 + t / 8, j0 + t / 8
// + 32, ... with one 16-byte load per row, so a workgroup keeps 32 rows in
// flight; the 32 row-lane sums are added in LDS in a fixed order.
__global__ __launch_bounds__(256) void k_pool_head(const int* __restrict__ seg, const short* __restrict__ H,
                                                   const float* __restrict__ w, const float* __restrict__ b,
                                                   float* __restrict__ pool, float* __restrict__ what) {
    const int k = blockIdx.x, fg = threadIdx.x & 7, rl = threadIdx.x >> 3;
    const int j0 = seg[k], j1 = seg[k + 1];
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int j = j0 + rl; j < j1; j += 32) {
        const s16x8 h = *reinterpret_cast<const s16x8*>(H + static_cast<size_t>(j) * kH + 8 * fg);
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] += bf2f(h[q]);
    }
    __shared__ float red[32][kH];
    __shared__ float dot[kH];
#pragma unroll
    for (int q = 0; q < 8; ++q) red[rl][8 * fg + q] = s[q];
    __syncthreads();
    if (threadIdx.x < kH) {
        const int c = threadIdx.x, cnt = j1 - j0;
        float t = 0.0f;
        for (int r = 0; r < 32; ++r) t += red[r][c];
        const float mean = cnt > 0 ? t / static_cast<float>(cnt) : 0.0f;
        pool[static_cast<size_t>(k) * kH + c] = mean;
        dot[c] = mean * w[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int e = 0; e < kH; ++e) t += dot[e];
        what[k] = t + b[0];
    }
}

// Loss term and parameter gradient of the primal-weight head (one
// workgroup), with weight LW = kOmegaLossWeight: dw[k] = 2 LW (what[k] -
// tw[k]) / count; loss += LW x mean of the squared errors; grad_w[c] =
// sum_k dw[k] pool[k][c]; grad_b = sum_k dw[k].  Adam scales each
// parameter's step, so the head learns at full speed while its small
// weight keeps it from pulling the shared layers away from x and y.
__global__ __launch_bounds__(kH) void k_omega_grad(int count, const float* __restrict__ what,
                                                   const float* __restrict__ tw, const float* __restrict__ pool,
                                                   float* __restrict__ dw, float* __restrict__ grad_w,
                                                   float* __restrict__ grad_b, float* __restrict__ loss) {
    const int c = threadIdx.x;
    for (int k = c; k < count; k += kH)
        dw[k] = 2.0f * kOmegaLossWeight * (what[k] - tw[k]) / static_cast<float>(count);
    __syncthreads();
    float g = 0.0f;
    for (int k = 0; k < count; ++k) g += dw[k] * pool[static_cast<size_t>(k) * kH + c];
    grad_w[c] = g;
    if (c == 0) {
        float gb = 0.0f, l = 0.0f;
        for (int k = 0; k < count; ++k) {
            gb += dw[k];
            const float e = what[k] - tw[k];
            l += e * e;
        }
        grad_b[0] = gb;
        loss[0] += kOmegaLossWeight * l / static_cast<float>(count);
    }
}

// dH[j][c] += dw[k] w[c] / (columns of k) for column j of segment k.
__global__ __launch_bounds__(256) void k_omega_bwd(int n, int count, const int* __restrict__ seg,
                                                   const float* __restrict__ dw, const float* __restrict__ w,
                                                   float* __restrict__ dH) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t j = t >> 6;
    const int c = static_cast<int>(t & 63);
    if (j >= n) return;
    int lo = 0, hi = count;  // the segment k with seg[k] <= j < seg[k+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (seg[mid] <= j) lo = mid;
        else hi = mid;
    }
    const int cnt = seg[lo + 1] - seg[lo];
    dH[j * kH + c] += dw[lo] * w[c] / static_cast<float>(cnt);
}

// Head: out[r] = H[r] . w + b.
__global__ __launch_bounds__(256) void k_head(int N, const short* __restrict__ Hin, const float* __restrict__ w,
                                              const float* __restrict__ b, float* __restrict__ out) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const short* h = Hin + static_cast<size_t>(r) * kH;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < kH / 8; ++q) {
        float v[8];
        load8_f(h + 8 * q, v);
#pragma unroll
        for (int e = 0; e < 8; ++e) s = __builtin_fmaf(v[e], w[8 * q + e], s);
    }
    out[r] = s + b[0];
}

// Loss residual: d[r] = 2 (pred - target) / N, and per-block sums of
// (pred - target)^2 / N into P[block].
__global__ __launch_bounds__(256) void k_mse(int N, const float* __restrict__ pred, const float* __restrict__ tgt,
                                             float* __restrict__ d, float* __restrict__ P) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    float e2 = 0.0f;
    if (r < N) {
        const float e = pred[r] - tgt[r];
        d[r] = 2.0f * e / static_cast<float>(N);
        e2 = e * e / static_cast<float>(N);
    }
    __shared__ float red[256];
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) P[blockIdx.x] = red[0];
}

// dH[r][c] = d[r] w[c]   (head backward, initialises dH)
__global__ __launch_bounds__(256) void k_head_bwd(int N, const float* __restrict__ d, const float* __restrict__ w,
                                                  float* __restrict__ dH) {
    const int t = blockIdx.x * 256 + threadIdx.x, r = t >> 6, c = t & 63;
    if (r >= N) return;
    dH[static_cast<size_t>(r) * kH + c] = d[r] * w[c];
}

// dP = dH * (Hout > 0), in place.
__global__ __launch_bounds__(256) void k_relu_bwd(size_t total, const short* __restrict__ Hout, float* __restrict__ dH) {
    const size_t t = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= total) return;
    if (!(bf2f(Hout[t]) > 0.0f)) dH[t] = 0.0f;
}

// Partial d-weighted row sums for the head weights: P[b][c] = sum_r d[r] H[r][c].
__global__ __launch_bounds__(256) void k_head_wgrad(int N, int rows_per, const float* __restrict__ d,
                                                    const short* __restrict__ Hin, float* __restrict__ P,
                                                    float* __restrict__ Pb) {
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int rb = blockIdx.x * rows_per;
    int re = rb + rows_per;
    if (re > N) re = N;
    float s = 0.0f, sb = 0.0f;
    for (int r = rb + g; r < re; r += 4) {
        s += d[r] * bf2f(Hin[static_cast<size_t>(r) * kH + c]);
        sb += d[r];
    }
    __shared__ float red[4][kH];
    __shared__ float redb[4][kH];
    red[g][c] = s;
    redb[g][c] = sb;
    __syncthreads();
    if (g == 0) {
        P[static_cast<size_t>(blockIdx.x) * kH + c] = red[0][c] + red[1][c] + red[2][c] + red[3][c];
        if (c == 0) Pb[blockIdx.x] = redb[0][0] + redb[1][0] + redb[2][0] + redb[3][0];
    }
}

__global__ __launch_bounds__(256) void k_adam(int np, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2, float lr, float c1,
                                              float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gi;
    const float b = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
}

// bfloat16 copies of a 64x64 weight: WT[c][k] = W[k][c] (forward) and
// WN[k][c] = W[k][c] (the B^T layout of W^T, for input gradients).
__global__ __launch_bounds__(256) void k_prep_w(const float* __restrict__ W, short* __restrict__ WT,
                                                short* __restrict__ WN) {
    for (int e = threadIdx.x; e < kH * kH; e += 256) {
        const int k = e / kH, c = e % kH;
        const short v = f2bf(W[e]);
        WT[c * kH + k] = v;
        WN[e] = v;
    }
}

// ---- parameter layout --------------------------------------------------
struct Offsets {
    int We_c, be_c, We_r, be_r;
    int Wrr[3], Wrz[3], br[3], Wcc[3], Wcz[3], bc[3];
    int wx, bx, wy, by, wom, bom, total;
    Offsets() {
        int o = 0;
        auto take = [&](int k) { const int r = o; o += k; return r; };
        We_c = take(kF * kH); be_c = take(kH); We_r = take(kF * kH); be_r = take(kH);
        for (int l = 1; l <= 2; ++l) {
            Wrr[l] = take(kH * kH); Wrz[l] = take(kH * kH); br[l] = take(kH);
            Wcc[l] = take(kH * kH); Wcz[l] = take(kH * kH); bc[l] = take(kH);
        }
        wx = take(kH); bx = take(1); wy = take(kH); by = take(1); wom = take(kH); bom = take(1);
        total = o;
    }
};
const Offsets kOff;

int cdiv(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace
}  // namespace pm

using namespace pm;

namespace pm {
namespace {

// Aggregation values in the other order, with the graph's checks: per row i
// (of nrows, CSR rp / ci over ncols), every index in range and strictly
// increasing, and each entry's value `vals` found at its place in the other
// orientation (cp / ri, rows ascending within a column, binary search) and
// written there in `out`: out[q] = vals[k] for the entry at CSR place k and
// the other orientation's place q.  Any failure sets *bad.
__global__ __launch_bounds__(256) void k_other_order(int nrows, int ncols, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const int* __restrict__ cp,
                                                     const int* __restrict__ ri, const float* __restrict__ vals,
                                                     float* __restrict__ out, int* __restrict__ bad) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= nrows) return;
    int prev = -1;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        if (j <= prev || j >= ncols) {
            bad[0] = 1;
            return;
        }
        prev = j;
        int lo = cp[j], hi = cp[j + 1];  // the row i among column j's rows
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ri[mid] < i) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= cp[j + 1] || ri[lo] != i) {
            bad[0] = 1;
            return;
        }
        out[lo] = vals[k];
    }
}

__global__ __launch_bounds__(256) void k_fill(int n, float* __restrict__ p, float v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// index of a dense 64x64 weight in the bf16 arrays: layer l (1,2), which:
// 0 Wrr, 1 Wrz, 2 Wcc, 3 Wcz
int widx(int l, int which) { return (l - 1) * 4 + which; }
int woff(int l, int which) {
    switch (which) {
        case 0: return kOff.Wrr[l];
        case 1: return kOff.Wrz[l];
        case 2: return kOff.Wcc[l];
        default: return kOff.Wcz[l];
    }
}

void fill(pm_gnn* g, float* p, int n, float v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(cdiv(n, 256)), dim3(256), 0, g->s, n, p, v);
    PM_LAUNCHED();
}

void prep_weights(pm_gnn* g) {
    for (int l = 1; l <= 2; ++l)
        for (int w = 0; w < 4; ++w) {
            hipLaunchKernelGGL(k_prep_w, dim3(1), dim3(256), 0, g->s, g->p.get() + woff(l, w),
                               g->wt.get() + widx(l, w) * kH * kH, g->wn.get() + widx(l, w) * kH * kH);
            PM_LAUNCHED();
        }
}

const short* WT(pm_gnn* g, int l, int w) { return g->wt.get() + widx(l, w) * kH * kH; }
const short* WN(pm_gnn* g, int l, int w) { return g->wn.get() + widx(l, w) * kH * kH; }

// Forward layer: out = relu(A1 B1 + A2 B2 + bias), bfloat16 in and out.
void dense(pm_gnn* g, int N, const short* A1, const short* B1, const short* A2, const short* B2, const float* bias,
           short* out) {
    if (N <= 0) return;
    hipLaunchKernelGGL((k_dense<true, true, true, false, short, short>), dim3(cdiv(N, 128)), dim3(256), 0, g->s, N,
                       A1, B1, A2, B2, bias, out);
    PM_LAUNCHED();
}

// Input gradient: out = D B, FP32 in and out.
void dense_grad(pm_gnn* g, int N, const float* D, const short* B, float* out) {
    if (N <= 0) return;
    hipLaunchKernelGGL((k_dense<false, false, false, false, float, float>), dim3(cdiv(N, 128)), dim3(256), 0, g->s,
                       N, D, B, nullptr, nullptr, nullptr, out);
    PM_LAUNCHED();
}

template <bool ACC, class IT, class OT>
void spmm(pm_gnn* g, int nrows, const int* rp, const int* idx, const float* val, const IT* Hin, OT* out) {
    if (nrows <= 0) return;
    const int per = cdiv(cdiv(nrows, 32), 8);
    hipLaunchKernelGGL((k_spmm<ACC, IT, OT>), dim3(8 * per), dim3(256), 0, g->s, nrows, per, rp, idx, val, Hin, out);
    PM_LAUNCHED();
}

constexpr int kRowsPer = 512;  // rows per workgroup in the gradient reductions

// out[0 .. width) = sum over nb partial rows of P (k_sum_partials).
void sum_partials(pm_gnn* g, int nb, int width, const float* P, float* out) {
    hipLaunchKernelGGL(k_sum_partials, dim3(cdiv(width, kSumLanes)), dim3(256), 0, g->s, nb, width, P, out);
    PM_LAUNCHED();
}

// grad[off .. off+4096) = A^T D summed over all N rows.
void wgrad(pm_gnn* g, int N, const short* A, const float* D, int off) {
    if (N <= 0) return fill(g, g->grad.get() + off, kH * kH, 0.0f);
    const int nb = cdiv(N, kRowsPer);
    hipLaunchKernelGGL(k_wgrad, dim3(nb), dim3(256), 0, g->s, N, kRowsPer, A, D, g->part.get());
    PM_LAUNCHED();
    sum_partials(g, nb, kH * kH, g->part.get(), g->grad.get() + off);
}

// grad[boff .. +64) = column sums of D; with X: grad[xoff .. +512) = X^T D.
void colsum(pm_gnn* g, int N, const float* D, const float* X, int boff, int xoff) {
    if (N <= 0) {
        fill(g, g->grad.get() + boff, kH, 0.0f);
        if (X) fill(g, g->grad.get() + xoff, kF * kH, 0.0f);
        return;
    }
    const int nb = cdiv(N, kRowsPer);
    float* P = g->part.get();
    float* P2 = P + static_cast<size_t>(nb) * kH;
    hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, g->s, N, kRowsPer, D, X, P, X ? P2 : nullptr);
    PM_LAUNCHED();
    sum_partials(g, nb, kH, P, g->grad.get() + boff);
    if (X) sum_partials(g, nb, kF * kH, P2, g->grad.get() + xoff);
}

void head_grads(pm_gnn* g, int N, const float* d, const short* Hin, int woff_, int boff) {
    if (N <= 0) {
        fill(g, g->grad.get() + woff_, kH, 0.0f);
        fill(g, g->grad.get() + boff, 1, 0.0f);
        return;
    }
    const int nb = cdiv(N, kRowsPer);
    float* P = g->part.get();
    float* Pb = P + static_cast<size_t>(nb) * kH;
    hipLaunchKernelGGL(k_head_wgrad, dim3(nb), dim3(256), 0, g->s, N, kRowsPer, d, Hin, P, Pb);
    PM_LAUNCHED();
    sum_partials(g, nb, kH, P, g->grad.get() + woff_);
    sum_partials(g, nb, 1, Pb, g->grad.get() + boff);
}

void relu_bwd(pm_gnn* g, int N, const short* Hout, float* dH) {
    if (N <= 0) return;
    const size_t total = static_cast<size_t>(N) * kH;
    hipLaunchKernelGGL(k_relu_bwd, dim3(cdiv(static_cast<long long>(total), 256)), dim3(256), 0, g->s, total, Hout, dH);
    PM_LAUNCHED();
}

void forward(pm_gnn* g) {
    const int m = g->m, n = g->n;
    const float* P = g->p.get();
    if (n > 0)
        hipLaunchKernelGGL(k_input, dim3(cdiv(static_cast<long long>(n) * 16, 256)), dim3(256), 0, g->s, n,
                           g->xc.get(), P + kOff.We_c, P + kOff.be_c, g->Hc[0].get());
    if (m > 0)
        hipLaunchKernelGGL(k_input, dim3(cdiv(static_cast<long long>(m) * 16, 256)), dim3(256), 0, g->s, m,
                           g->xr.get(), P + kOff.We_r, P + kOff.be_r, g->Hr[0].get());
    PM_LAUNCHED();
    for (int l = 1; l <= 2; ++l) {
        spmm<false>(g, m, g->rp.get(), g->ci.get(), g->ar.get(), g->Hc[l - 1].get(), g->Zr[l].get());
        dense(g, m, g->Hr[l - 1].get(), WT(g, l, 0), g->Zr[l].get(), WT(g, l, 1), P + kOff.br[l], g->Hr[l].get());
        spmm<false>(g, n, g->cp.get(), g->ri.get(), g->ac.get(), g->Hr[l].get(), g->Zc[l].get());
        dense(g, n, g->Hc[l - 1].get(), WT(g, l, 2), g->Zc[l].get(), WT(g, l, 3), P + kOff.bc[l], g->Hc[l].get());
    }
    if (n > 0)
        hipLaunchKernelGGL(k_head, dim3(cdiv(n, 256)), dim3(256), 0, g->s, n, g->Hc[2].get(), P + kOff.wx, P + kOff.bx,
                           g->xhat.get());
    if (m > 0)
        hipLaunchKernelGGL(k_head, dim3(cdiv(m, 256)), dim3(256), 0, g->s, m, g->Hr[2].get(), P + kOff.wy, P + kOff.by,
                           g->yhat.get());
    hipLaunchKernelGGL(k_pool_head, dim3(g->nseg), dim3(256), 0, g->s, g->seg.get(), g->Hc[2].get(), P + kOff.wom,
                       P + kOff.bom, g->pool.get(), g->what.get());
    PM_LAUNCHED();
}

// Loss and gradient for targets already in g->tx / g->ty.
float backward(pm_gnn* g) {
    const int m = g->m, n = g->n;
    const float* P = g->p.get();
    float* part = g->part.get();
    const int nbx = cdiv(std::max(n, 1), 256), nby = cdiv(std::max(m, 1), 256);
    // MSE residuals; loss partials at the end of `part` (kept apart from
    // the gradient partials by summing them first).
    if (n > 0) hipLaunchKernelGGL(k_mse, dim3(nbx), dim3(256), 0, g->s, n, g->xhat.get(), g->tx.get(), g->dx.get(), part);
    if (m > 0)
        hipLaunchKernelGGL(k_mse, dim3(nby), dim3(256), 0, g->s, m, g->yhat.get(), g->ty.get(), g->dy.get(),
                           part + (n > 0 ? nbx : 0));
    PM_LAUNCHED();
    float* loss_dev = g->grad.get() + kOff.total;  // one spare slot after the gradient
    sum_partials(g, (n > 0 ? nbx : 0) + (m > 0 ? nby : 0), 1, part, loss_dev);
    head_grads(g, n, g->dx.get(), g->Hc[2].get(), kOff.wx, kOff.bx);
    head_grads(g, m, g->dy.get(), g->Hr[2].get(), kOff.wy, kOff.by);
    float* dHc = g->dHc.get();
    float* dHr = g->dHr.get();
    float* dHc2 = g->dHc2.get();
    float* dZ = g->dZ.get();
    if (n > 0) {
        hipLaunchKernelGGL(k_head_bwd, dim3(cdiv(static_cast<long long>(n) * kH, 256)), dim3(256), 0, g->s, n,
                           g->dx.get(), P + kOff.wx, dHc);
        PM_LAUNCHED();
    }
    if (g->has_tw) {
        hipLaunchKernelGGL(k_omega_grad, dim3(1), dim3(kH), 0, g->s, g->nseg, g->what.get(), g->tw.get(),
                           g->pool.get(), g->dw.get(), g->grad.get() + kOff.wom, g->grad.get() + kOff.bom, loss_dev);
        if (n > 0)
            hipLaunchKernelGGL(k_omega_bwd, dim3(cdiv(static_cast<long long>(n) * kH, 256)), dim3(256), 0, g->s, n,
                               g->nseg, g->seg.get(), g->dw.get(), P + kOff.wom, dHc);
        PM_LAUNCHED();
    } else {
        fill(g, g->grad.get() + kOff.wom, kH + 1, 0.0f);
    }
    if (m > 0) {
        hipLaunchKernelGGL(k_head_bwd, dim3(cdiv(static_cast<long long>(m) * kH, 256)), dim3(256), 0, g->s, m,
                           g->dy.get(), P + kOff.wy, dHr);
        PM_LAUNCHED();
    }
    for (int l = 2; l >= 1; --l) {
        // column update of layer l: dHc holds dL/dHc(l)
        relu_bwd(g, n, g->Hc[l].get(), dHc);
        wgrad(g, n, g->Hc[l - 1].get(), dHc, kOff.Wcc[l]);
        wgrad(g, n, g->Zc[l].get(), dHc, kOff.Wcz[l]);
        colsum(g, n, dHc, nullptr, kOff.bc[l], 0);
        dense_grad(g, n, dHc, WN(g, l, 2), dHc2);                                       // dHc(l-1)
        dense_grad(g, n, dHc, WN(g, l, 3), dZ);                                         // dZc
        spmm<true>(g, m, g->rp.get(), g->ci.get(), g->ac_t.get(), dZ, dHr);             // dHr += Ac^T dZc
        // row update of layer l: dHr holds dL/dHr(l)
        relu_bwd(g, m, g->Hr[l].get(), dHr);
        wgrad(g, m, g->Hr[l - 1].get(), dHr, kOff.Wrr[l]);
        wgrad(g, m, g->Zr[l].get(), dHr, kOff.Wrz[l]);
        colsum(g, m, dHr, nullptr, kOff.br[l], 0);
        dense_grad(g, m, dHr, WN(g, l, 1), dZ);                                         // dZr
        dense_grad(g, m, dHr, WN(g, l, 0), g->dT.get());                                // dHr(l-1)
        spmm<true>(g, n, g->cp.get(), g->ri.get(), g->ar_t.get(), dZ, dHc2);            // dHc(l-1) += Ar^T dZr
        std::swap(g->dHc, g->dHc2);
        std::swap(g->dHr, g->dT);
        dHc = g->dHc.get();
        dHc2 = g->dHc2.get();
        dHr = g->dHr.get();
    }
    relu_bwd(g, n, g->Hc[0].get(), dHc);
    colsum(g, n, dHc, g->xc.get(), kOff.be_c, kOff.We_c);
    relu_bwd(g, m, g->Hr[0].get(), dHr);
    colsum(g, m, dHr, g->xr.get(), kOff.be_r, kOff.We_r);
    float loss = 0.0f;
    PM_HIP(hipMemcpyAsync(&loss, loss_dev, sizeof(float), hipMemcpyDeviceToHost, g->s));
    PM_HIP(hipStreamSynchronize(g->s));
    return loss;
}

}  // namespace
}  // namespace pm

namespace pm {

void gnn_reserve(pm_gnn* g, int m, int n, int nnz, bool training) {
    g->m = m;
    g->n = n;
    g->nnz = nnz;
    const size_t mr = std::max(m, 1), nc = std::max(n, 1);
    for (int l = 0; l < 3; ++l) {
        g->Hc[l].resize(nc * kH);
        g->Hr[l].resize(mr * kH);
        g->Zr[l].resize(mr * kH);
        g->Zc[l].resize(nc * kH);
    }
    g->xhat.resize(nc);
    g->yhat.resize(mr);
    const int whole[2] = {0, n};  // one segment until pm_gnn_set_segments says otherwise
    gnn_set_segments(g, 1, whole);
    g->has_tw = false;
    g->grad_ready = training;
    if (!training) return;
    g->dHc.resize(nc * kH);
    g->dHc2.resize(nc * kH);
    g->dHr.resize(mr * kH);
    g->dZ.resize(std::max(mr, nc) * kH);
    g->dT.resize(mr * kH);
    for (auto* b : {&g->dx, &g->tx}) b->resize(nc);
    for (auto* b : {&g->dy, &g->ty}) b->resize(mr);
    const size_t nb = static_cast<size_t>(cdiv(std::max(m, n), kRowsPer));
    g->part.resize(std::max(nb * kH * kH, static_cast<size_t>(cdiv(std::max(m, 1), 256) + cdiv(std::max(n, 1), 256))) +
                   nb * (kF + 1) * kH + 64);
}

void gnn_forward(pm_gnn* g) { forward(g); }

void gnn_set_segments(pm_gnn* g, int count, const int* col_off) {
    if (count < 1 || !col_off) fail(PM_ERR_INVALID_ARG, "segments: need at least one");
    if (col_off[0] != 0 || col_off[count] != g->n) fail(PM_ERR_INVALID_ARG, "segments must run from 0 to n");
    for (int k = 0; k < count; ++k)
        if (col_off[k + 1] < col_off[k]) fail(PM_ERR_INVALID_ARG, "segment offsets decrease");
    g->nseg = count;
    g->seg.upload(col_off, static_cast<size_t>(count) + 1, g->s);
    g->pool.resize(static_cast<size_t>(count) * kH);
    for (auto* b : {&g->what, &g->tw, &g->dw}) b->resize(count);
    g->has_tw = false;
}

}  // namespace pm

extern "C" {

int32_t pm_gnn_num_params(void) { return pm::kOff.total; }

int pm_gnn_create(int32_t device, uint64_t seed, pm_gnn** out) {
    return pm::guarded([&] {
        using namespace pm;
        if (!out) fail(PM_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        int nd = 0;
        if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
            (void)hipGetLastError();
            fail(PM_ERR_NO_DEVICE, "no HIP device visible");
        }
        if (device < 0 || device >= nd) fail(PM_ERR_NO_DEVICE, "device ordinal out of range");
        PM_HIP(hipSetDevice(device));
        auto g = std::make_unique<pm_gnn>();
        g->device = device;
        PM_HIP(hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking));
        // Scaled uniform init U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from SplitMix64.
        std::vector<float> h(kOff.total, 0.0f);
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + 1;
        auto uni = [&]() {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            t ^= t >> 31;
            return static_cast<double>(t >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        };
        auto init = [&](int off, int count, int fan_in) {
            const double a = 1.0 / std::sqrt(static_cast<double>(fan_in));
            for (int i = 0; i < count; ++i) h[off + i] = static_cast<float>(a * uni());
        };
        init(kOff.We_c, kF * kH, kF);
        init(kOff.We_r, kF * kH, kF);
        for (int l = 1; l <= 2; ++l) {
            init(kOff.Wrr[l], kH * kH, 2 * kH);
            init(kOff.Wrz[l], kH * kH, 2 * kH);
            init(kOff.Wcc[l], kH * kH, 2 * kH);
            init(kOff.Wcz[l], kH * kH, 2 * kH);
        }
        init(kOff.wx, kH, kH);
        init(kOff.wy, kH, kH);
        g->p.upload(h, g->s);
        g->grad.resize(kOff.total + 1);
        g->m1.resize(kOff.total);
        g->m2.resize(kOff.total);
        fill(g.get(), g->m1.get(), kOff.total, 0.0f);
        fill(g.get(), g->m2.get(), kOff.total, 0.0f);
        g->wt.resize(8 * kH * kH);
        g->wn.resize(8 * kH * kH);
        prep_weights(g.get());
        PM_HIP(hipStreamSynchronize(g->s));
        *out = g.release();
    });
}

int pm_gnn_destroy(pm_gnn* g) {
    return pm::guarded([&] {
        if (!g) return;
        (void)hipSetDevice(g->device);
        if (g->s) (void)hipStreamSynchronize(g->s);
        hipStream_t s = g->s;
        delete g;
        if (s) (void)hipStreamDestroy(s);
    });
}

int pm_gnn_get_params(pm_gnn* g, float* out) {
    return pm::guarded([&] {
        if (!g || !out) pm::fail(PM_ERR_INVALID_ARG, "NULL argument");
        PM_HIP(hipSetDevice(g->device));
        g->p.download(out, pm::kOff.total, g->s);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_gnn_set_params(pm_gnn* g, const float* in) {
    return pm::guarded([&] {
        using namespace pm;
        if (!g || !in) fail(PM_ERR_INVALID_ARG, "NULL argument");
        PM_HIP(hipSetDevice(g->device));
        g->p.upload(in, kOff.total, g->s);
        fill(g, g->m1.get(), kOff.total, 0.0f);
        fill(g, g->m2.get(), kOff.total, 0.0f);
        g->step = 0;
        prep_weights(g);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_host_alloc(size_t bytes, void** out) {
    return pm::guarded([&] {
        if (!out) pm::fail(PM_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        PM_HIP(hipHostMalloc(out, bytes > 0 ? bytes : 1, hipHostMallocDefault));
    });
}

int pm_host_free(void* p) {
    return pm::guarded([&] {
        if (p) PM_HIP(hipHostFree(p));
    });
}

int pm_gnn_set_graph(pm_gnn* g, int32_t m, int32_t n, int32_t nnz, const int32_t* rp, const int32_t* ci,
                     const float* ar, const int32_t* cp, const int32_t* ri, const float* ac, const float* xr,
                     const float* xc) {
    return pm::guarded([&] {
        using namespace pm;
        if (!g || m < 0 || n < 0 || nnz < 0 || !rp || !cp) fail(PM_ERR_INVALID_ARG, "bad graph");
        if (nnz > 0 && (!ci || !ar || !ri || !ac)) fail(PM_ERR_INVALID_ARG, "NULL graph arrays");
        if ((m > 0 && !xr) || (n > 0 && !xc)) fail(PM_ERR_INVALID_ARG, "NULL features");
        if (rp[0] != 0 || rp[m] != nnz || cp[0] != 0 || cp[n] != nnz) fail(PM_ERR_INVALID_ARG, "CSR inconsistent");
        for (int i = 0; i < m; ++i)
            if (rp[i] > rp[i + 1]) fail(PM_ERR_INVALID_ARG, "rp decreasing");
        for (int j = 0; j < n; ++j)
            if (cp[j] > cp[j + 1]) fail(PM_ERR_INVALID_ARG, "cp decreasing");
        PM_HIP(hipSetDevice(g->device));
        hipStream_t s = g->s;
        g->rp.upload(rp, static_cast<size_t>(m) + 1, s);
        g->cp.upload(cp, static_cast<size_t>(n) + 1, s);
        g->ci.upload(ci, nnz, s);
        g->ri.upload(ri, nnz, s);
        g->ar.upload(ar, nnz, s);
        g->ac.upload(ac, nnz, s);
        g->xr.upload(xr, static_cast<size_t>(m) * kF, s);
        g->xc.upload(xc, static_cast<size_t>(n) * kF, s);
        // Transposed-order copies of the aggregation values for the backward
        // pass, made on the device with the index checks (indices in range,
        // strictly increasing within each row and column, both orientations
        // the same entries): ar_t (Ar^T, CSR over columns), ac_t (Ac^T, CSR
        // over rows).
        g->ar_t.resize(std::max(nnz, 1));
        g->ac_t.resize(std::max(nnz, 1));
        g->flag.resize(1);
        PM_HIP(hipMemsetAsync(g->flag.get(), 0, sizeof(int), s));
        if (m > 0 && nnz > 0) {
            hipLaunchKernelGGL(k_other_order, dim3((m + 255) / 256), dim3(256), 0, s, m, n, g->rp.get(), g->ci.get(),
                               g->cp.get(), g->ri.get(), g->ar.get(), g->ar_t.get(), g->flag.get());
            hipLaunchKernelGGL(k_other_order, dim3((n + 255) / 256), dim3(256), 0, s, n, m, g->cp.get(), g->ri.get(),
                               g->rp.get(), g->ci.get(), g->ac.get(), g->ac_t.get(), g->flag.get());
            PM_LAUNCHED();
        }
        int bad = 0;
        g->flag.download(&bad, 1, s);
        PM_HIP(hipStreamSynchronize(s));
        if (bad) {
            g->m = g->n = g->nnz = 0;
            fail(PM_ERR_INVALID_ARG, "graph indices out of range, not increasing within a row or column, or the "
                                     "two orientations disagree (the graph is cleared)");
        }
        g->m = m;
        g->n = n;
        g->nnz = nnz;
        pm::gnn_reserve(g, m, n, nnz, true);
        PM_HIP(hipStreamSynchronize(s));
    });
}

int pm_gnn_set_segments(pm_gnn* g, int32_t count, const int32_t* col_off) {
    return pm::guarded([&] {
        if (!g) pm::fail(PM_ERR_INVALID_ARG, "NULL network");
        PM_HIP(hipSetDevice(g->device));
        pm::gnn_set_segments(g, count, col_off);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_gnn_set_omega_targets(pm_gnn* g, const float* targets) {
    return pm::guarded([&] {
        if (!g) pm::fail(PM_ERR_INVALID_ARG, "NULL network");
        PM_HIP(hipSetDevice(g->device));
        g->has_tw = targets != nullptr;
        if (targets) g->tw.upload(targets, g->nseg, g->s);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_gnn_forward_omega(pm_gnn* g, float* omega_hat) {
    return pm::guarded([&] {
        if (!g || !omega_hat) pm::fail(PM_ERR_INVALID_ARG, "NULL argument");
        PM_HIP(hipSetDevice(g->device));
        pm::forward(g);
        g->what.download(omega_hat, g->nseg, g->s);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_gnn_set_omega_head(pm_gnn* g, int32_t enabled) {
    return pm::guarded([&] {
        if (!g) pm::fail(PM_ERR_INVALID_ARG, "NULL network");
        g->omega_head = enabled != 0;
    });
}

int pm_gnn_forward(pm_gnn* g, float* xhat, float* yhat) {
    return pm::guarded([&] {
        if (!g) pm::fail(PM_ERR_INVALID_ARG, "NULL network");
        PM_HIP(hipSetDevice(g->device));
        pm::forward(g);
        if (xhat && g->n) g->xhat.download(xhat, g->n, g->s);
        if (yhat && g->m) g->yhat.download(yhat, g->m, g->s);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_gnn_gradient(pm_gnn* g, const float* tx, const float* ty, float* loss, float* grad) {
    return pm::guarded([&] {
        if (!g || !tx || !ty) pm::fail(PM_ERR_INVALID_ARG, "NULL argument");
        if (!g->grad_ready) pm::fail(PM_ERR_INVALID_ARG, "no training graph: call pm_gnn_set_graph first");
        PM_HIP(hipSetDevice(g->device));
        if (g->n) g->tx.upload(tx, g->n, g->s);
        if (g->m) g->ty.upload(ty, g->m, g->s);
        pm::forward(g);
        const float l = pm::backward(g);
        if (loss) *loss = l;
        if (grad) g->grad.download(grad, pm::kOff.total, g->s);
        PM_HIP(hipStreamSynchronize(g->s));
    });
}

int pm_gnn_train_step(pm_gnn* g, const float* tx, const float* ty, float lr, float* loss) {
    return pm::guarded([&] {
        using namespace pm;
        if (!g || !tx || !ty) fail(PM_ERR_INVALID_ARG, "NULL argument");
        if (!g->grad_ready) pm::fail(PM_ERR_INVALID_ARG, "no training graph: call pm_gnn_set_graph first");
        PM_HIP(hipSetDevice(g->device));
        if (g->n) g->tx.upload(tx, g->n, g->s);
        if (g->m) g->ty.upload(ty, g->m, g->s);
        forward(g);
        const float l = backward(g);
        ++g->step;
        const float c1 = 1.0f - std::pow(0.9f, static_cast<float>(g->step));
        const float c2 = 1.0f - std::pow(0.999f, static_cast<float>(g->step));
        hipLaunchKernelGGL(k_adam, dim3(cdiv(kOff.total, 256)), dim3(256), 0, g->s, kOff.total, g->p.get(),
                           g->grad.get(), g->m1.get(), g->m2.get(), lr, c1, c2);
        PM_LAUNCHED();
        prep_weights(g);
        PM_HIP(hipStreamSynchronize(g->s));
        if (loss) *loss = l;
    });
}

}  // extern "C"
