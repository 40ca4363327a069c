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
// pm_ruiz_scale: the host arrays go to the device once, the passes run
// there, and the scaled arrays come back.
// pm_ruiz_scale_noalloc: the same passes on arrays already in device
// memory, with the per-pass factors and column norms in this file's
// static device array `scratch` instead of allocated buffers; its kernels
// check every row pointer and column index against the sizes given and
// report a bad one instead of reading outside the arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
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

// Scratch of pm_ruiz_scale_noalloc: word 0 is its error flag (1 = a row
// pointer or column index outside the sizes given), then, from word 32
// (256 bytes), the row factors [m], column factors [n] and column norms
// [n] of a pass.
constexpr int64_t kScratchHead = 32;
constexpr int64_t kNoallocMaxEntries = int64_t{1} << 22;  // m + 2 n at most
static __device__ __align__(256) double scratch[kScratchHead + kNoallocMaxEntries];

// Sets the error flag (a plain store of 1: any order of writers gives 1).
__device__ inline void flag_bad(unsigned long long* err) { *err = 1ull; }

__global__ __launch_bounds__(1) void k_clear_flag(unsigned long long* err) { *err = 0ull; }

// cn[col] <- max(cn[col], |v_k|) over every nonzero.  GUARD: column indexes
// outside [0, n) are flagged and skipped.
template <bool GUARD>
__global__ __launch_bounds__(kBlock) void k_col_norm(int64_t nnz, int n, const int* __restrict__ ci,
                                                     const double* __restrict__ v, Bits* __restrict__ cn,
                                                     unsigned long long* err) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; k < nnz; k += stride) {
        const Bits a = static_cast<Bits>(__double_as_longlong(fabs(v[k])));
        const int j = ci[k];
        if (GUARD && (j < 0 || j >= n)) {
            flag_bad(err);
            continue;
        }
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

// Entry range of row i; GUARD: a range outside [0, nnz) or decreasing is
// flagged and treated as empty.
template <bool GUARD>
__device__ inline int2 row_range(const int* __restrict__ rp, int i, int64_t nnz, unsigned long long* err) {
    int b = rp[i], e = rp[i + 1];
    if (GUARD && (b < 0 || e < b || static_cast<int64_t>(e) > nnz)) {
        flag_bad(err);
        b = e = 0;
    }
    return make_int2(b, e);
}

// rs[i] <- 1/sqrt(max_k |v_k|) over row i (1 for a zero row), L lanes per
// row.
template <int L, bool GUARD>
__global__ __launch_bounds__(kBlock) void k_row_norm_l(int m, int64_t nnz, const int* __restrict__ rp,
                                                       const double* __restrict__ v, double* __restrict__ rs,
                                                       unsigned long long* err) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t / L, lane = t % L;
    const bool on = i < m;
    double a = 0.0;
    if (on) {
        const int2 r = row_range<GUARD>(rp, i, nnz, err);
        for (int k = r.x + lane; k < r.y; k += L) a = fmax(a, fabs(v[k]));
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, L));
    if (on && lane == 0) rs[i] = a > 0.0 ? 1.0 / sqrt(a) : 1.0;
}

// A's rows (v *= rs * cs[col]), L lanes per row; the finite row bounds and
// row_scale by the row's first lane.  GUARD: bad ranges and column
// indexes are flagged and skipped.
template <int L, bool GUARD>
__global__ __launch_bounds__(kBlock) void k_row_apply(int m, int n, int64_t nnz, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, double* __restrict__ v,
                                                      const double* __restrict__ rs, const double* __restrict__ cs,
                                                      double* __restrict__ lo, double* __restrict__ hi,
                                                      double* __restrict__ acc, unsigned long long* err) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t / L, lane = t % L;
    if (i >= m) return;
    const double r = rs[i];
    const int2 rr = row_range<GUARD>(rp, i, nnz, err);
    for (int k = rr.x + lane; k < rr.y; k += L) {
        const int j = ci[k];
        if (GUARD && (j < 0 || j >= n)) {
            flag_bad(err);
            continue;
        }
        v[k] *= r * cs[j];
    }
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

// The arrays of one scaling run on the device.
struct RuizArrays {
    int m, n;
    int64_t nnz;
    const int *rp, *ci;
    double *v, *c, *rl, *ru, *vl, *vu;  // scaled in place
    double *dr, *dc;                     // cumulative factors (outputs)
    double *rs, *cs;                     // factors of a pass
    Bits* cn;                            // column norms of a pass
    unsigned long long* err;             // GUARD runs only
};

template <int L, bool GUARD>
void row_pass(hipStream_t s, const RuizArrays& a, bool norm) {
    const dim3 g(static_cast<unsigned>((static_cast<int64_t>(a.m) * L + kBlock - 1) / kBlock)), b(kBlock);
    if (norm)
        hipLaunchKernelGGL((k_row_norm_l<L, GUARD>), g, b, 0, s, a.m, a.nnz, a.rp, a.v, a.rs, a.err);
    else
        hipLaunchKernelGGL((k_row_apply<L, GUARD>), g, b, 0, s, a.m, a.n, a.nnz, a.rp, a.ci, a.v, a.rs, a.cs, a.rl,
                           a.ru, a.dr, a.err);
}

template <bool GUARD>
void rows(int lanes, hipStream_t s, const RuizArrays& a, bool norm) {
    switch (lanes) {
        case 1: return row_pass<1, GUARD>(s, a, norm);
        case 2: return row_pass<2, GUARD>(s, a, norm);
        case 4: return row_pass<4, GUARD>(s, a, norm);
        case 8: return row_pass<8, GUARD>(s, a, norm);
        case 16: return row_pass<16, GUARD>(s, a, norm);
        case 32: return row_pass<32, GUARD>(s, a, norm);
        default: return row_pass<64, GUARD>(s, a, norm);
    }
}

// `passes` Ruiz passes over arrays already on the device (factors set to 1
// first); launches only, on stream s.
template <bool GUARD>
void ruiz_passes(hipStream_t s, const RuizArrays& a, int passes) {
    const dim3 b(kBlock);
    hipLaunchKernelGGL(k_ones, grid_of(a.m), b, 0, s, a.m, a.dr);
    hipLaunchKernelGGL(k_ones, grid_of(a.n), b, 0, s, a.n, a.dc);
    // one nonzero per thread per step of a grid-stride loop
    const int64_t nz_blocks = a.nnz > 0 ? std::min<int64_t>((a.nnz + kBlock - 1) / kBlock, 16384) : 0;
    const int lanes = row_lanes(static_cast<double>(a.nnz) / a.m);
    for (int32_t pass = 0; pass < passes; ++pass) {
        rows<GUARD>(lanes, s, a, true);
        hipLaunchKernelGGL(k_col_norm_init, grid_of(a.n), b, 0, s, a.n, a.c, a.cn);
        if (nz_blocks > 0)
            hipLaunchKernelGGL(k_col_norm<GUARD>, dim3(static_cast<unsigned>(nz_blocks)), b, 0, s, a.nnz, a.n, a.ci,
                               a.v, a.cn, a.err);
        hipLaunchKernelGGL(k_col_apply, grid_of(a.n), b, 0, s, a.n, a.cn, a.cs, a.c, a.vl, a.vu, a.dc);
        rows<GUARD>(lanes, s, a, false);
        PM_LAUNCHED();
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
    const RuizArrays a{m, n, nnz, d_rp.get(), d_ci.get(), d_v.get(), d_c.get(), d_rl.get(), d_ru.get(), d_vl.get(),
                       d_vu.get(), d_dr.get(), d_dc.get(), d_rs.get(), d_cs.get(), d_cn.get(), nullptr};
    ruiz_passes<false>(s, a, passes);
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

// Address of `scratch` on each device, looked up once per device (each GPU
// holds its own copy of the array).  Host code uses these addresses, never
// the array's name, which on the host is only a stand-in.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

double* scratch_of_current_device(int& dev) {
    PM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(PM_ERR_INVALID_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        PM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(PM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

void ruiz_noalloc(int32_t m, int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci, double* v, double* c,
                  double* rl, double* ru, double* vl, double* vu, int32_t passes, double* row_scale,
                  double* col_scale, hipStream_t s) {
    if (m < 0 || n < 0 || nnz < 0) fail(PM_ERR_INVALID_ARG, "negative dimension");
    if (passes < 0) fail(PM_ERR_INVALID_ARG, "negative number of passes");
    if (m > 0 && (!rp || !rl || !ru || !row_scale)) fail(PM_ERR_INVALID_ARG, "row_ptr / row bounds / row_scale is NULL");
    if (n > 0 && (!c || !vl || !vu || !col_scale)) fail(PM_ERR_INVALID_ARG, "c / variable bounds / col_scale is NULL");
    if (nnz > 0 && (!ci || !v)) fail(PM_ERR_INVALID_ARG, "col_idx / values is NULL");
    if (nnz > INT32_MAX) fail(PM_ERR_INVALID_ARG, "more than 2^31 - 1 nonzeros");
    const int64_t need = static_cast<int64_t>(m) + 2 * static_cast<int64_t>(n);
    if (need > kNoallocMaxEntries)
        fail(PM_ERR_INVALID_ARG, "m + 2 n = " + std::to_string(need) + " exceeds the scratch of pm_ruiz_scale_noalloc (" +
                                     std::to_string(kNoallocMaxEntries) + " values)");
    int dev = 0;
    double* base = scratch_of_current_device(dev);
    auto* err = reinterpret_cast<unsigned long long*>(base);
    const dim3 b(kBlock);
    if (m > 0) hipLaunchKernelGGL(k_ones, grid_of(m), b, 0, s, m, row_scale);
    if (n > 0) hipLaunchKernelGGL(k_ones, grid_of(n), b, 0, s, n, col_scale);
    PM_LAUNCHED();
    if (m == 0 || n == 0 || passes == 0) {  // scaling.cpp:37: nothing to scale
        PM_HIP(hipStreamSynchronize(s));
        return;
    }
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, err);
    double* rs = base + kScratchHead;
    const RuizArrays a{m, n, nnz, rp, ci, v, c, rl, ru, vl, vu, row_scale, col_scale, rs, rs + m,
                       reinterpret_cast<Bits*>(rs + m + n), err};
    ruiz_passes<true>(s, a, passes);
    unsigned long long bad = 0;  // into host stack memory: no allocation
    PM_HIP(hipMemcpyAsync(&bad, err, sizeof(bad), hipMemcpyDeviceToHost, s));
    PM_HIP(hipStreamSynchronize(s));
    if (bad) fail(PM_ERR_INVALID_ARG, "row_ptr or col_idx outside the sizes given (the arrays are partly scaled)");
}

}  // namespace
}  // namespace pm

extern "C" {

int pm_ruiz_scale_noalloc(int32_t m, int32_t n, int32_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                          double* values, double* c, double* row_lower, double* row_upper, double* var_lower,
                          double* var_upper, int32_t passes, double* row_scale, double* col_scale, void* stream) {
    return pm::guarded([&] {
        pm::ruiz_noalloc(m, n, nnz, row_ptr, col_idx, values, c, row_lower, row_upper, var_lower, var_upper, passes,
                         row_scale, col_scale, static_cast<hipStream_t>(stream));
    });
}

int pm_ruiz_scale(int32_t m, int32_t n, const int32_t* row_ptr, const int32_t* col_idx, double* values, double* c,
                  double* row_lower, double* row_upper, double* var_lower, double* var_upper, int32_t passes,
                  double* row_scale, double* col_scale) {
    pm::Range range_("pm_ruiz_scale");
    return pm::guarded([&] {
        pm::check_inputs(m, n, row_ptr, col_idx, values, c, row_lower, row_upper, var_lower, var_upper, passes,
                         row_scale, col_scale);
        pm::ruiz_device(m, n, row_ptr, col_idx, values, c, row_lower, row_upper, var_lower, var_upper, passes,
                        row_scale, col_scale);
    });
}

}  // extern "C"
