// batch.hip -- many independent LPs solved together by one replayed graph.
//
// Layout: every LP's A and A^T (CSR, local indices) and vectors are packed
// back to back in shared device arrays, i.e. the batch is one block-diagonal
// LP; a per-LP descriptor holds the offsets, and global row / column pointer
// arrays address the concatenation.  A sliced ELL copy of A and A^T (rows
// sorted by length within groups of 1024 rows of A and 2048 of A^T, 64-row
// slices, 16-bit indexes relative to each group's column window) feeds the
// sliced iterate kernels.
//
// Creation: host threads transpose each LP and build its layouts; the
// device scales the whole concatenation (Ruiz + Pock-Chambolle act on one
// row or column at a time, so each LP gets its own factors); the autotuner
// times every iterate kernel on the full batch and keeps the fastest.
//
// Phases:
//  init -- one 1024-thread workgroup per LP (largest first): power
//    iteration for ||A_hat||_2 (first run only), the initial primal weight,
//    the judgement of a warm start by its FP64 KKT error.
//  iterate -- the hot loop: each reflected-Halpern iteration is two
//    grid-wide launches over a work list of the running LPs' row groups
//    (A^T y -> x, A xe -> y) with per-LP step sizes, Halpern weights and
//    chunk lengths carried by each item.  Each XCD owns one contiguous
//    eighth of the list (its LPs' vectors stay in its L2).  Grids are sized
//    to the kernel's occupancy.  Once few LPs are left a second graph
//    switches to lanes-per-row kernels with many small items per LP.
//  round / check -- grid-wide passes over 128-row blocks of every running
//    LP: a refinement round's correction-LP data, the total iterates, the
//    FP64 relative-KKT terms reduced per block; then one workgroup per LP
//    sums its blocks, takes the pm::Control decision (the code the single-LP
//    host loop runs) and applies the restart or refinement it asks for; the
//    work list is rebuilt.
// Per-LP solver state lives in device memory between phases.  A few
// round/iterate/check chunks are captured once in a HIP graph; the host
// replays it until a device counter reports no active LP.  Nothing is
// allocated or cleared with a memset between replays.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <type_traits>
#include <vector>

#include "batch_internal.hpp"
#include "control.hpp"
#include "dispatch.hpp"
#include "prep.hpp"
#include "scaling.hpp"

namespace pm {
namespace {

// Workgroup size of the batch kernel: 16 wavefronts per LP, so one LP keeps
// enough loads in flight to stream its data from HBM.
constexpr int kBB = 1024;

// Rows per block of the round and check passes, and doubles per block of
// the check's partial sums.
constexpr int kCheckRows = 128, kPart = 8;

// Largest lane count the iterate kernels are built for.
constexpr int kMaxLanes = 32;
// Nonzeros of the running LPs below which the batch switches to the tail
// kernels (lanes per row by mean row length): about where a sliced launch,
// a few workgroups per LP, stops being throughput-bound.
constexpr int64_t kTailNnz = int64_t(3) << 20;

// Kernel of an explicit SpMV variant: lanes per row, kStreamKernel,
// kSlicedKernel, a staged code, or kAutoKernel for PM_SPMV_AUTO and the variants the
// iterate kernels do not offer (autotuned instead).
constexpr int kAutoKernel = -100;
inline int variant_kernel(int v) {
    switch (v) {
        case PM_SPMV_SCALAR: return 1;
        case PM_SPMV_VECTOR2: return 2;
        case PM_SPMV_VECTOR4: return 4;
        case PM_SPMV_VECTOR8: return 8;
        case PM_SPMV_VECTOR16: return 16;
        case PM_SPMV_VECTOR32: return 32;
        case PM_SPMV_STREAM: return kStreamKernel;
        case PM_SPMV_SLICED: return kSlicedKernel;
        case PM_SPMV_STAGED: return kStagedKernel;
        default: return kAutoKernel;
    }
}

inline int lanes_variant(int l) {
    switch (l) {
        case kStreamKernel: return PM_SPMV_STREAM;
        case kSlicedKernel: return PM_SPMV_SLICED;
        case kStagedKernel: return PM_SPMV_STAGED;
        case 1: return PM_SPMV_SCALAR;
        case 2: return PM_SPMV_VECTOR2;
        case 4: return PM_SPMV_VECTOR4;
        case 16: return PM_SPMV_VECTOR16;
        case 32: return PM_SPMV_VECTOR32;
        default: return PM_SPMV_VECTOR8;
    }
}

// Sliced ELL form of one CSR (local indices), appended to the host arrays:
// groups of `group` rows, rows sorted by length (longest first, ties in
// row order) inside each group, slices of kSliceRows sorted rows; each
// row's entries stay in CSR order.
struct SlicedHost {
    std::vector<int> src;  // entry of the CSR value array each slot takes, -1 for padding
    std::vector<uint16_t> idx;
    std::vector<int> len, pos;
    std::vector<int64_t> ptr;
    std::vector<int2> win;
};

// Padding entries hold value 0 and the row's last column (the window's
// first column for an empty row), so that they read an element the group
// reads anyway.  The values are filled on the device from the scaled CSR
// values: slot e takes value[src[e]] (global entry indexes from nz0).
inline void build_sliced(int nrows, const int* rp, const int* ci, int64_t nz0, SlicedHost& out, int group) {
    std::vector<int> order;
    for (int g0 = 0; g0 < nrows; g0 += group) {
        const int gn = std::min(group, nrows - g0);
        int lo = INT_MAX, hi = -1;
        for (int k = rp[g0]; k < rp[g0 + gn]; ++k) {
            lo = std::min(lo, ci[k]);
            hi = std::max(hi, ci[k]);
        }
        if (hi < 0) lo = 0, hi = -1;
        const int2 win = make_int2(lo, hi - lo + 1);
        order.resize(gn);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return rp[g0 + a + 1] - rp[g0 + a] > rp[g0 + b + 1] - rp[g0 + b];
        });
        for (int q = 0; q * kSliceRows < gn; ++q) {
            int len = 0;
            for (int l = 0; l < kSliceRows && q * kSliceRows + l < gn; ++l) {
                const int r = g0 + order[q * kSliceRows + l];
                len = std::max(len, rp[r + 1] - rp[r]);
            }
            out.ptr.push_back(static_cast<int64_t>(out.src.size()));
            out.len.push_back(len);
            out.win.push_back(win);
            for (int l = 0; l < kSliceRows; ++l) {
                const int p = q * kSliceRows + l;
                out.pos.push_back(p < gn ? order[p] : -1);
            }
            const size_t base = out.src.size();
            out.src.resize(base + static_cast<size_t>(len) * kSliceRows, -1);
            out.idx.resize(base + static_cast<size_t>(len) * kSliceRows, 0);
            for (int l = 0; l < kSliceRows; ++l) {
                const int p = q * kSliceRows + l;
                if (p >= gn) continue;
                const int r = g0 + order[p];
                for (int k = rp[r]; k < rp[r + 1]; ++k) {
                    const size_t e = base + static_cast<size_t>(k - rp[r]) * kSliceRows + l;
                    out.src[e] = static_cast<int>(nz0 + k);
                    out.idx[e] = static_cast<uint16_t>(ci[k] - lo);  // exact when the window fits
                }
                const int last = rp[r + 1] > rp[r] ? ci[rp[r + 1] - 1] - lo : 0;
                for (int j = rp[r + 1] - rp[r]; j < len; ++j)
                    out.idx[base + static_cast<size_t>(j) * kSliceRows + l] = static_cast<uint16_t>(last);
            }
        }
    }
}

// Rows per sliced group of an LP's A: kSliceGroupWide when its groups of
// kSliceGroup rows span on average more than half of its n columns (random
// column structure: the window stays near n whatever the group) and n fits
// the staged kernel's LDS window, kSliceGroup otherwise (banded structure,
// such as multi-period plans, whose windows grow with the group).
inline int group_rows_a(int m, int n, const int* rp, const int* ci) {
    if (m <= kSliceGroup || n > kStagedMaxVec) return kSliceGroup;
    double width = 0.0;
    int groups = 0;
    for (int g0 = 0; g0 < m; g0 += kSliceGroup, ++groups) {
        const int g1 = std::min(m, g0 + kSliceGroup);
        int lo = INT_MAX, hi = -1;
        for (int k = rp[g0]; k < rp[g1]; ++k) {
            lo = std::min(lo, ci[k]);
            hi = std::max(hi, ci[k]);
        }
        if (hi >= lo) width += hi - lo + 1;
    }
    return width > 0.5 * static_cast<double>(n) * groups ? kSliceGroupWide : kSliceGroup;
}

// Rows per group when `rows` rows are cut into groups of at most `group`
// rows of about equal size (a multiple of the slice height): the fewest
// groups whose last one keeps at least half a group.
inline int balanced_group(int rows, int group) {
    if (rows <= group) return group;
    for (int count = (rows + group - 1) / group;; ++count) {
        const int g = ((rows + count - 1) / count + kSliceRows - 1) / kSliceRows * kSliceRows;
        if (g <= kSliceRows || 2 * (rows - (count - 1) * g) >= g) return g;
    }
}

// Size class of an LP for the per-class iterate launches: fewer than 256
// rows, fewer than 1,024 (one group of A), or more.
inline int size_class(int m) { return m < 256 ? 0 : (m < kSliceGroup ? 1 : 2); }

// Stream segments of one CSR: consecutive rows with at most kStreamNnz
// nonzeros and kBlock rows together; a longer row forms a segment alone.
inline void stream_segments(int nrows, const int* rp, std::vector<int2>& seg) {
    seg.clear();
    int first = 0;
    while (first < nrows) {
        int last = first + 1;
        while (last < nrows && last - first < kBlock && rp[last + 1] - rp[first] <= kStreamNnz) ++last;
        seg.push_back(make_int2(first, last - first));
        first = last;
    }
}

template <class VT>
struct BatchArrays {
    const int* rp;
    const int* ci;
    const float* va;
    const VT* oa;
    const int* cp;
    const int* ri;
    const float* vat;
    const VT* oat;
    const double *cmin, *lr, *ur, *lx, *ux, *dr, *dc;
    const double *wx, *wy;  // warm-start base points (original space)
    float *c32, *lr32, *ur32, *lx32, *ux32, *g32;
    float *x, *xe, *x0, *xt, *xl, *y, *y0, *yt, *yl;
    double *xtot, *ytot, *xbase, *ybase;
    double *xsave, *ysave, *xpol;  // feasibility polishing: saved candidate, polished primal
    float* hold32;                 // the main solve's state during a polish (pm_batch::hold32)
    double* hold64;
    int64_t cols_total, rows_total;
    BatchDesc* desc;
    const int* order;
    const int* grp;     // global row pointer of A   (nz offsets of the concatenation)
    const int* gcp;     // global row pointer of A^T
    LpIter* it;
    int4* wrange;
    int2 *ck_r, *ck_c;
    double *part_r, *part_c;
    const int* wcnt;
    // sliced A (sa_*) and A^T (st_*)
    const float *sa_val, *st_val;
    const uint16_t *sa_idx, *st_idx;
    const int *sa_ptr, *sa_len, *sa_pos, *st_ptr, *st_len, *st_pos;
    const int2 *sa_win, *st_win;
    double tol;
    int max_iterations, check_frequency, max_refinements, polish;
};

// ---- block helpers -------------------------------------------------------

template <int NS, int NM, int NT = kBB>
__device__ void block_reduce_all(double (&acc)[NS + NM], double* out) {
    __shared__ double tmp[NT / 64][16];
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
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 0; j < NS + NM; ++j) tmp[threadIdx.x >> 6][j] = acc[j];
    }
    __syncthreads();
    if (threadIdx.x < NS + NM) {
        const int j = threadIdx.x;
        double a = tmp[0][j];
        for (int w = 1; w < NT / 64; ++w) a = (j < NS) ? a + tmp[w][j] : (tmp[w][j] > a ? tmp[w][j] : a);
        out[j] = a;
    }
    __syncthreads();
}

template <int L, class AT, class VT, class XT, class Epi>
__device__ __forceinline__ void brows(int nrows, const int* rp, const int* ci, const VT* v, const XT* x,
                                      const Epi& epi) {
    constexpr int G = kBB / L;
    const int g = threadIdx.x / L, lane = threadIdx.x & (L - 1);
    for (int r0 = 0; r0 < nrows; r0 += G) {
        const int row = r0 + g;
        if (row < nrows) {
            const AT s = row_dot<L, false, AT>(rp[row], rp[row + 1], lane, ci, v, x);
            if (lane == 0) epi(row, s);
        }
    }
}

template <class AT, class VT, class XT, class Epi>
__device__ void brows_any(int L, int nrows, const int* rp, const int* ci, const VT* v, const XT* x,
                          const Epi& epi) {
    switch (L) {
        case 1: brows<1, AT>(nrows, rp, ci, v, x, epi); break;
        case 2: brows<2, AT>(nrows, rp, ci, v, x, epi); break;
        case 4: brows<4, AT>(nrows, rp, ci, v, x, epi); break;
        case 8: brows<8, AT>(nrows, rp, ci, v, x, epi); break;
        case 16: brows<16, AT>(nrows, rp, ci, v, x, epi); break;
        default: brows<32, AT>(nrows, rp, ci, v, x, epi); break;
    }
}

// Reduction row pass: per-row terms (assigned by the epilogue into t[])
// accumulated per thread (sums, maxima), then reduced over the block into
// out[] (shared, visible to all threads on return).
template <int L, class VT, class XT, class Epi>
__device__ __forceinline__ void brows_red(int nrows, const int* rp, const int* ci, const VT* v, const XT* x,
                                          const Epi& epi, double (&acc)[Epi::NS + Epi::NM]) {
    constexpr int G = kBB / L;
    const int g = threadIdx.x / L, lane = threadIdx.x & (L - 1);
    for (int r0 = 0; r0 < nrows; r0 += G) {
        const int row = r0 + g;
        if (row < nrows) {
            const double s = row_dot<L, false, double>(rp[row], rp[row + 1], lane, ci, v, x);
            if (lane == 0) {
                double t[Epi::NS + Epi::NM];
                epi(row, s, t);
#pragma unroll
                for (int j = 0; j < Epi::NS + Epi::NM; ++j)
                    acc[j] = j < Epi::NS ? acc[j] + t[j] : (t[j] > acc[j] ? t[j] : acc[j]);
            }
        }
    }
}

template <class VT, class XT, class Epi>
__device__ __attribute__((noinline)) void brows_red_any(int L, int nrows, const int* rp, const int* ci, const VT* v, const XT* x,
                              const Epi& epi, double* out) {
    double acc[Epi::NS + Epi::NM];
#pragma unroll
    for (int j = 0; j < Epi::NS + Epi::NM; ++j) acc[j] = 0.0;
    switch (L) {
        case 1: brows_red<1>(nrows, rp, ci, v, x, epi, acc); break;
        case 2: brows_red<2>(nrows, rp, ci, v, x, epi, acc); break;
        case 4: brows_red<4>(nrows, rp, ci, v, x, epi, acc); break;
        case 8: brows_red<8>(nrows, rp, ci, v, x, epi, acc); break;
        case 16: brows_red<16>(nrows, rp, ci, v, x, epi, acc); break;
        default: brows_red<32>(nrows, rp, ci, v, x, epi, acc); break;
    }
    block_reduce_all<Epi::NS, Epi::NM>(acc, out);
}

// Per-LP pointers.
template <class VT>
struct View {
    int m, n, la, lat;
    const int *rp, *ci, *cp, *ri;
    const float *va, *vat;
    const VT *oa, *oat;
    const double *cmin, *lr, *ur, *lx, *ux, *dr, *dc, *wx, *wy;
    float *c32, *lr32, *ur32, *lx32, *ux32, *g32, *x, *xe, *x0, *xt, *xl, *y, *y0, *yt, *yl;
    double *xtot, *ytot, *xbase, *ybase, *xsave, *ysave, *xpol;
    float *hx[6], *hy[6];  // held main state: x, x0, xl, c32, lx32, ux32 / y, y0, yl, lr32, ur32, g32
    double *hxb, *hyb;     // held x_base, y_base
};

template <class VT>
__device__ __forceinline__ View<VT> view_of(const BatchArrays<VT>& A, const BatchDesc& d) {
    View<VT> v;
    v.m = d.m;
    v.n = d.n;
    v.la = d.lanes_a;
    v.lat = d.lanes_at;
    v.rp = A.rp + d.rp_off;
    v.cp = A.cp + d.cp_off;
    v.ci = A.ci + d.nz_off;
    v.ri = A.ri + d.nz_off;
    v.va = A.va + d.nz_off;
    v.vat = A.vat + d.nz_off;
    v.oa = A.oa + d.nz_off;
    v.oat = A.oat + d.nz_off;
    const int64_t ro = d.row_off, co = d.col_off;
    v.cmin = A.cmin + co; v.lx = A.lx + co; v.ux = A.ux + co; v.dc = A.dc + co; v.wx = A.wx + co;
    v.lr = A.lr + ro; v.ur = A.ur + ro; v.dr = A.dr + ro; v.wy = A.wy + ro;
    v.c32 = A.c32 + co; v.lx32 = A.lx32 + co; v.ux32 = A.ux32 + co;
    v.x = A.x + co; v.xe = A.xe + co; v.x0 = A.x0 + co; v.xt = A.xt + co; v.xl = A.xl + co;
    v.xtot = A.xtot + co; v.xbase = A.xbase + co; v.xsave = A.xsave + co; v.xpol = A.xpol + co;
    v.lr32 = A.lr32 + ro; v.ur32 = A.ur32 + ro; v.g32 = A.g32 + ro;
    v.y = A.y + ro; v.y0 = A.y0 + ro; v.yt = A.yt + ro; v.yl = A.yl + ro;
    v.ytot = A.ytot + ro; v.ybase = A.ybase + ro; v.ysave = A.ysave + ro;
    for (int q = 0; q < 6; ++q) {
        v.hx[q] = A.hold32 + q * A.cols_total + co;
        v.hy[q] = A.hold32 + 6 * A.cols_total + q * A.rows_total + ro;
    }
    v.hxb = A.hold64 + co;
    v.hyb = A.hold64 + A.cols_total + ro;
    return v;
}

// Solver state of one LP between the phase kernels.
struct BatchState {
    Control ctl;
    Control ctl_held;  // the main solve's control state while a polish runs
    Polish pol;
    double eta;
    int32_t status;
    int32_t active;       // 1 while the LP is being solved
    int32_t round_start;  // 1: the next round-data phase must run
    int32_t have_cand;    // a candidate T(z) exists in xtot / ytot
};

// The next chunk of LP lp as the iterate kernels read it.
__device__ __forceinline__ void publish_iter(LpIter* it, int lp, const BatchState& st, const BatchDesc& d,
                                             int K) {
    LpIter r;
    const double om = st.ctl.omega;
    r.tau = static_cast<float>(st.eta / om);
    r.sig = static_cast<float>(st.eta * om);
    r.k0 = st.pol.phase != Polish::MAIN ? st.pol.k : st.ctl.k;
    r.chunk = st.active ? st.ctl.chunk(K) : 0;
    // A polished pair rides one iteration (whose iterates are discarded)
    // so that the grid-wide passes evaluate its KKT error.
    if (st.pol.phase == Polish::VERIFY && r.chunk > 1) r.chunk = 1;
    r.use_g = st.ctl.use_g ? 1 : 0;
    r.row_off = static_cast<int32_t>(d.row_off);
    r.col_off = static_cast<int32_t>(d.col_off);
    r.pad = 0;
    it[lp] = r;
}

// FP64 relative KKT of (xtot, ytot) into the descriptor; returns rel on
// thread 0 (every thread must call).
// Relative KKT error of a total iterate from the reduced row terms kr
// (EpiKktRow) and column terms kc (EpiKktCol); stored into d.
__device__ inline double kkt_finish(const double* kr, const double* kc, BatchDesc& d) {
    const double p = kc[4], dd2 = kr[5] + kc[3];
    const double rp = sqrt(kr[0]) / (1.0 + sqrt(kr[2]));
    const double rdd = sqrt(kr[3] + kc[0]) / (1.0 + sqrt(kc[2]));
    const double rg = fabs(p - dd2) / (1.0 + fabs(p) + fabs(dd2));
    double rel = rp > rdd ? rp : rdd;
    rel = rel > rg ? rel : rg;
    if (!(rel == rel)) rel = INFINITY;
    d.rel_kkt = rel; d.rel_p = rp; d.rel_d = rdd; d.rel_g = rg; d.pobj = p; d.dobj = dd2;
    return rel;
}

template <class VT>
__device__ double batch_kkt(const View<VT>& v, BatchDesc& d, double* red, double* keep) {
    brows_red_any(v.la, v.m, v.rp, v.ci, v.oa, v.xtot, EpiKktRow{v.lr, v.ur, v.ytot, v.dr}, red);
    if (threadIdx.x < 7) keep[threadIdx.x] = red[threadIdx.x];
    __syncthreads();
    brows_red_any(v.lat, v.n, v.cp, v.ri, v.oat, v.ytot, EpiKktCol{v.cmin, v.lx, v.ux, v.xtot, v.dc}, red);
    double rel = 0.0;
    if (threadIdx.x == 0) rel = kkt_finish(keep, red, d);
    __syncthreads();
    return rel;
}

// Phase 0: power iteration for ||A_hat||_2 (first run of the batch only),
// initial primal weight, base points, and the judgement of a warm start.
template <class VT>
__global__ __launch_bounds__(kBB) void k_b_init(BatchArrays<VT> A, BatchState* S) {
    const int lp = A.order[blockIdx.x];
    BatchDesc& d = A.desc[lp];
    const View<VT> v = view_of(A, d);
    const int m = v.m, n = v.n, tid = threadIdx.x;
    __shared__ double red[16], keep[8];
    __shared__ double rel_s[2];
    const uint64_t t0 = wall_clock64();
    double lam = 0.0;
    if (d.eta <= 0.0) {  // ||A_hat||_2 by power iteration, once per batch
        // At least 48 steps, then until ||A^T A v||^2 changes by <= 1e-5
        // relative, at most 1,024: the estimate approaches sigma_max from
        // below, and a fixed count stopped 0.8 % short on a window whose
        // top singular values cluster (0.8571, 0.8567, 0.8566).
        for (int j = tid; j < n; j += kBB) v.xt[j] = rsqrtf(static_cast<float>(n));
        __syncthreads();
        if (m > 0 && d.nnz > 0) {
            double prev = -1.0;
            for (int it = 0; it < 1024; ++it) {
                brows_any<float>(v.la, m, v.rp, v.ci, v.va, v.xt, EpiStore<float>{v.yt});
                __syncthreads();
                brows_any<float>(v.lat, n, v.cp, v.ri, v.vat, v.yt, EpiStore<float>{v.xe});
                __syncthreads();
                double acc[1] = {0.0};
                for (int j = tid; j < n; j += kBB) acc[0] += static_cast<double>(v.xe[j]) * v.xe[j];
                block_reduce_all<1, 0>(acc, red);
                lam = red[0];
                const double inv = lam > 0.0 ? 1.0 / sqrt(lam) : 0.0;
                for (int j = tid; j < n; j += kBB) v.xt[j] = static_cast<float>(v.xe[j] * inv);
                __syncthreads();
                // every thread holds the same lam (block_reduce_all), so all leave together
                if (it + 1 >= 48 && (!(lam > 0.0) || fabs(lam - prev) <= 1e-5 * lam)) break;
                prev = lam;
            }
        }
    }
    double acc[2] = {0.0, 0.0};
    for (int j = tid; j < n; j += kBB) {
        const double t = v.dc[j] * v.cmin[j];
        acc[0] += t * t;
        double xb = d.warm_x ? v.wx[j] : 0.0;
        if (d.warm_x) xb = xb < v.lx[j] ? v.lx[j] : (xb > v.ux[j] ? v.ux[j] : xb);
        v.xbase[j] = xb;
    }
    for (int i = tid; i < m; i += kBB) {
        double bh = 0.0;
        if (v.lr[i] > -kInf) bh = fabs(v.lr[i]);
        if (v.ur[i] < kInf && fabs(v.ur[i]) > bh) bh = fabs(v.ur[i]);
        const double t = v.dr[i] * bh;
        acc[1] += t * t;
        double yb = d.warm_y ? v.wy[i] : 0.0;
        if (!(v.ur[i] < kInf) && yb > 0.0) yb = 0.0;  // the sign each row allows
        if (!(v.lr[i] > -kInf) && yb < 0.0) yb = 0.0;
        v.ybase[i] = yb;
    }
    block_reduce_all<2, 0>(acc, red);
    const double c_norm2 = red[0], b_norm2 = red[1];
    // A warm start is judged by its FP64 relative KKT error against the
    // cold start's: the worse of the two is not used.
    bool warm = d.warm_x || d.warm_y;
    if (warm) {
        for (int j = tid; j < n; j += kBB) v.xtot[j] = v.xbase[j];
        for (int i = tid; i < m; i += kBB) v.ytot[i] = v.ybase[i];
        __syncthreads();
        const double rw = batch_kkt(v, d, red, keep);
        if (tid == 0) rel_s[0] = rw;
        for (int j = tid; j < n; j += kBB) v.xtot[j] = v.lx[j] > 0.0 ? v.lx[j] : (v.ux[j] < 0.0 ? v.ux[j] : 0.0);
        for (int i = tid; i < m; i += kBB) v.ytot[i] = 0.0;
        __syncthreads();
        const double r0 = batch_kkt(v, d, red, keep);
        if (tid == 0) rel_s[1] = r0;
        __syncthreads();
        warm = rel_s[0] <= rel_s[1];
        if (!warm) {
            for (int j = tid; j < n; j += kBB) v.xbase[j] = 0.0;
            for (int i = tid; i < m; i += kBB) v.ybase[i] = 0.0;
        }
        __syncthreads();
    }
    if (tid == 0) {
        BatchState& st = S[lp];
        double om = (c_norm2 > 1e-20 && b_norm2 > 1e-20) ? sqrt(c_norm2) / sqrt(b_norm2) : 1.0;
        if (d.omega0 > 0.0) om = d.omega0;
        else if (d.omega_scale > 0.0 && warm) om *= d.omega_scale;
        if (d.eta <= 0.0) d.eta = lam > 0.0 ? 0.99 / sqrt(sqrt(lam)) : 1.0;
        d.warm_accepted = (d.warm_x || d.warm_y) ? (warm ? 1 : 0) : -1;
        d.start_kkt = (d.warm_x || d.warm_y) ? (warm ? rel_s[0] : rel_s[1]) : -1.0;
        st.ctl.init(A.tol, A.max_iterations, A.max_refinements, om, warm && d.warm_y != 0 && m > 0);
        st.pol.init(A.polish != 0);
        st.eta = d.eta;
        st.status = PM_SOLVE_ITERATION_LIMIT;
        st.active = 1;
        st.round_start = 1;
        st.have_cand = 0;
        d.t_start = t0;
        publish_iter(A.it, lp, st, d, A.check_frequency);
    }
}


// Work list of the iterate kernels, rebuilt after every check: one
// WorkItem per group of G rows of every LP still running (G = 256 / lanes;
// G = 0 selects the LP's stream segments, G < 0 the groups of the sliced
// layout), in LP order, carrying that LP's parameters for the chunk;
// counts in wcnt[0] (rows of A) and wcnt[1] (rows of A^T).  Also the
// blocks of the round and check passes (ck_r, ck_c, counts in wcnt[2],
// wcnt[3], each LP's range in wrange).  One 1024-thread workgroup scans
// the LPs.  bucket: -1 both lists over every LP; -2 the check blocks only;
// c >= 0 the items of the LPs of size class c only (wr, wc and wcnt then
// point at that class's regions; wcnt[2], wcnt[3] are left alone).
template <class VT>
__global__ __launch_bounds__(1024) void k_b_worklist(BatchArrays<VT> A, int count, int bucket, int ga, int gt,
                                                     const int2* __restrict__ seg_a, const int2* __restrict__ seg_at,
                                                     WorkItem* __restrict__ wr, WorkItem* __restrict__ wc,
                                                     int* __restrict__ wcnt) {
    __shared__ int sr[1024], sc[1024], kr[1024], kc[1024], fa[1024], act[1024];
    __shared__ int carry_r, carry_c, carry_kr, carry_kc;
    const int tid = static_cast<int>(threadIdx.x);
    // ga, gt: rows per item of the lanes kernels, 0 for stream segments,
    // negative for groups of the sliced layout (each LP's own group rows,
    // BatchDesc::ga / gt).
    if (tid == 0) carry_r = carry_c = carry_kr = carry_kc = 0;
    __syncthreads();
    for (int base = 0; base < count; base += 1024) {
        const int k = base + tid;
        int nr = 0, nc = 0, m = 0, n = 0, br = 0, bc = 0;
        LpIter P{};
        if (k < count) P = A.it[k];
        if (k < count && P.chunk > 0) {
            const BatchDesc& d = A.desc[k];
            m = d.m;
            n = d.n;
            if (bucket == -1 || d.bucket == bucket) {
                nr = ga > 0 ? (m + ga - 1) / ga : (ga == 0 ? d.sa_cnt : (m + d.ga - 1) / d.ga);
                nc = gt > 0 ? (n + gt - 1) / gt : (gt == 0 ? d.st_cnt : (n + d.gt - 1) / d.gt);
            }
            if (bucket < 0) {
                br = (m + kCheckRows - 1) / kCheckRows;
                bc = (n + kCheckRows - 1) / kCheckRows;
            }
        }
        sr[tid] = nr;
        sc[tid] = nc;
        kr[tid] = br;
        kc[tid] = bc;
        fa[tid] = nr + nc + br + bc > 0 ? 1 : 0;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // inclusive scans (Hillis-Steele)
            const int ar = tid >= o ? sr[tid - o] : 0, ac = tid >= o ? sc[tid - o] : 0;
            const int akr = tid >= o ? kr[tid - o] : 0, akc = tid >= o ? kc[tid - o] : 0;
            const int af = tid >= o ? fa[tid - o] : 0;
            __syncthreads();
            sr[tid] += ar;
            sc[tid] += ac;
            kr[tid] += akr;
            kc[tid] += akc;
            fa[tid] += af;
            __syncthreads();
        }
        if (k < count && bucket < 0)
            A.wrange[k] = make_int4(carry_kr + kr[tid] - br, br, carry_kc + kc[tid] - bc, bc);
        // The LPs of this chunk with anything to write, compacted; each
        // gets `per` threads, which write its items and blocks interleaved
        // (a few LPs left running spread over the whole workgroup).
        if (nr + nc + br + bc > 0) act[fa[tid] - 1] = tid;
        __syncthreads();
        const int na = fa[1023];
        const int per = na > 0 ? max(1, 1024 / na) : 1024;
        const int slot = tid / per, sub = tid - slot * per;
        if (slot < na) {
            const int q = act[slot], kk = base + q;
            const int nrk = sr[q] - (q ? sr[q - 1] : 0), nck = sc[q] - (q ? sc[q - 1] : 0);
            const int brk = kr[q] - (q ? kr[q - 1] : 0), bck = kc[q] - (q ? kc[q - 1] : 0);
            const int offr = carry_r + sr[q] - nrk, offc = carry_c + sc[q] - nck;
            const int offkr = carry_kr + kr[q] - brk, offkc = carry_kc + kc[q] - bck;
            const LpIter Pk = A.it[kk];
            const BatchDesc& d = A.desc[kk];
            const int mk = d.m, nk = d.n;
            const int rok = static_cast<int>(d.row_off), cok = static_cast<int>(d.col_off);
            for (int g = sub; g < brk; g += per) A.ck_r[offkr + g] = make_int2(kk, g);
            for (int g = sub; g < bck; g += per) A.ck_c[offkc + g] = make_int2(kk, g);
            // The last field carries the LP, or for the sliced and staged
            // kernels the item's first slice.
            for (int g = sub; g < nrk; g += per) {
                int first = g * ga, rows = min(ga, mk - g * ga), tag = kk;
                if (ga == 0) {
                    const int2 sg = seg_a[d.sa_off + g];
                    first = sg.x;
                    rows = sg.y;
                } else if (ga < 0) {
                    first = g * d.ga;
                    rows = min(d.ga, mk - first);
                    tag = d.la_off + first / kSliceRows;
                }
                wr[offr + g] = WorkItem{rok + first, rows, Pk.k0, Pk.chunk, Pk.use_g, Pk.col_off, Pk.sig, tag};
            }
            for (int g = sub; g < nck; g += per) {
                int first = g * gt, rows = min(gt, nk - g * gt), tag = kk;
                if (gt == 0) {
                    const int2 sg = seg_at[d.st_off + g];
                    first = sg.x;
                    rows = sg.y;
                } else if (gt < 0) {
                    first = g * d.gt;
                    rows = min(d.gt, nk - first);
                    tag = d.lt_off + first / kSliceRows;
                }
                wc[offc + g] = WorkItem{cok + first, rows, Pk.k0, Pk.chunk, 0, Pk.row_off, Pk.tau, tag};
            }
        }
        __syncthreads();
        if (tid == 1023) {
            carry_r += sr[1023];
            carry_c += sc[1023];
            carry_kr += kr[1023];
            carry_kc += kc[1023];
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (bucket != -2) {
            wcnt[0] = carry_r;
            wcnt[1] = carry_c;
        }
        if (bucket < 0) {
            wcnt[2] = carry_kr;
            wcnt[3] = carry_kc;
        }
    }
}

// Phase 2: iteration i of the chunk, both halves as grid-wide launches over
// the work list.  Block b runs on XCD b % 8 and takes the items of that
// XCD's eighth of the list (contiguous LPs, whose vectors then stay in the
// XCD's L2), strided by the number of blocks per XCD.
__device__ __forceinline__ void xcd_slice(int total, int& w0, int& w1, int& step) {
    const int nx = gridDim.x >= 8 ? 8 : 1;
    const int x = static_cast<int>(blockIdx.x) % nx;
    w0 = static_cast<int>(static_cast<int64_t>(total) * x / nx) + static_cast<int>(blockIdx.x) / nx;
    w1 = static_cast<int>(static_cast<int64_t>(total) * (x + 1) / nx);
    step = static_cast<int>(gridDim.x) / nx;
}

// The next item is read while the current one is computed, so the only
// dependent loads left per item are its row pointers, matrix entries and
// gathered vector entries.  Items are read as two int4 (wave-uniform, so
// scalar loads) and kept in registers.
struct ItemRegs {
    int4 a, b;  // a = {row0, cnt, k0, chunk}, b = {use_g, off, step bits, lp}
};

__device__ __forceinline__ ItemRegs load_item(const WorkItem* work, int w) {
    const int4* p = reinterpret_cast<const int4*>(work + w);
    return ItemRegs{p[0], p[1]};
}

template <int L, class VT>
__global__ __launch_bounds__(kBlock) void k_s_primal(BatchArrays<VT> A, const WorkItem* __restrict__ work,
                                                     const int* __restrict__ wcnt, int i) {
    static_assert(sizeof(WorkItem) == 32, "WorkItem is read as two int4");
    const int lane = static_cast<int>(threadIdx.x) & (L - 1), g = static_cast<int>(threadIdx.x) / L;
    int w0, w1, step;
    xcd_slice(wcnt[1], w0, w1, step);
    if (w0 >= w1) return;
    ItemRegs cur = load_item(work, w0);
    for (int w = w0; w < w1; w += step) {
        ItemRegs nxt = cur;
        if (w + step < w1) nxt = load_item(work, w + step);
        const int chunk = cur.a.w;
        if (g < cur.a.y && i < chunk) {
            const int j = cur.a.x + g;
            const bool keep = i == chunk - 1;
            const EpiPrimalHal epi{__int_as_float(cur.b.z), Control::lambda(cur.a.z + i), 1.0f, A.c32, A.lx32,
                                   A.ux32, A.x0, A.x, A.xe, keep ? A.xt : nullptr};
            EpiPrimalHal::Pre pre{};
            if (lane == 0) pre = epi.load(j);  // in flight with the product's loads
            const float aty = row_dot<L, false, float>(A.gcp[j], A.gcp[j + 1], lane, A.ri, A.vat, A.y + cur.b.y);
            if (lane == 0) epi.apply(j, aty, pre);
        }
        cur = nxt;
    }
}

template <int L, class VT>
__global__ __launch_bounds__(kBlock) void k_s_dual(BatchArrays<VT> A, const WorkItem* __restrict__ work,
                                                   const int* __restrict__ wcnt, int i) {
    const int lane = static_cast<int>(threadIdx.x) & (L - 1), g = static_cast<int>(threadIdx.x) / L;
    int w0, w1, step;
    xcd_slice(wcnt[0], w0, w1, step);
    if (w0 >= w1) return;
    ItemRegs cur = load_item(work, w0);
    for (int w = w0; w < w1; w += step) {
        ItemRegs nxt = cur;
        if (w + step < w1) nxt = load_item(work, w + step);
        const int chunk = cur.a.w;
        if (g < cur.a.y && i < chunk) {
            const int r = cur.a.x + g;
            const bool keep = i == chunk - 1;
            const EpiDualHal epi{__int_as_float(cur.b.z), Control::lambda(cur.a.z + i), 1.0f, A.lr32, A.ur32,
                                 cur.b.x ? A.g32 : nullptr, A.y0, A.y, keep ? A.yt : nullptr};
            EpiDualHal::Pre pre{};
            if (lane == 0) pre = epi.load(r);  // in flight with the product's loads
            const float ax = row_dot<L, false, float>(A.grp[r], A.grp[r + 1], lane, A.ci, A.va, A.xe + cur.b.y);
            if (lane == 0) epi.apply(r, ax, pre);
        }
        cur = nxt;
    }
}

// Stream variant: a work item is a segment of consecutive rows holding at
// most kStreamNnz nonzeros (a longer row alone), split over the
// workgroup's four wavefronts.  Each wavefront reads its rows' nonzeros as
// one contiguous range, 64 consecutive entries per load (every load fully
// coalesced, unlike L-lanes-per-row, where a wavefront touches a line per
// row per step), stages the products in its own slice of LDS, and one lane
// per row sums that row in CSR order.  No workgroup barrier: wavefronts
// synchronise only among their own lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool PRIMAL, class VT>
__global__ __launch_bounds__(kBlock) void k_t_half(BatchArrays<VT> A, const WorkItem* __restrict__ work,
                                                   const int* __restrict__ wcnt, int i) {
    constexpr int kWaves = kBlock / 64, kWaveBuf = 512;
    __shared__ float buf[kWaves][kWaveBuf];
    const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
    float* lb = buf[wv];
    const int* rp = PRIMAL ? A.gcp : A.grp;
    const int* idx = PRIMAL ? A.ri : A.ci;
    const float* val = PRIMAL ? A.vat : A.va;
    int w0, w1, step;
    xcd_slice(wcnt[PRIMAL ? 1 : 0], w0, w1, step);
    for (int w = w0; w < w1; w += step) {
        const ItemRegs it = load_item(work, w);
        const int chunk = it.a.w;
        if (i >= chunk) continue;
        const int per = (it.a.y + kWaves - 1) / kWaves;  // rows per wavefront, <= 64
        const int r0 = it.a.x + wv * per, r1 = min(it.a.x + it.a.y, r0 + per);
        if (r0 >= r1) continue;  // uniform over the wavefront
        const float* xv = (PRIMAL ? A.y : A.xe) + it.b.y;
        const int nz0 = rp[r0], nz1 = rp[r1];
        const bool has = lane < r1 - r0;
        int rb = 0, re = 0;
        if (has) {
            rb = rp[r0 + lane];
            re = rp[r0 + lane + 1];
        }
        float s = 0.0f;
        for (int base = nz0; base < nz1; base += kWaveBuf) {
            const int top = min(nz1, base + kWaveBuf);
            for (int k = base + lane; k < top; k += 64) lb[k - base] = val[k] * xv[idx[k]];
            wave_sync();
            if (has) {
                const int a = max(rb, base), e = min(re, top);
                for (int k = a; k < e; ++k) s += lb[k - base];
            }
            wave_sync();
        }
        if (has) {
            const int r = r0 + lane;
            const bool keep = i == chunk - 1;
            const float stp = __int_as_float(it.b.z), lam = Control::lambda(it.a.z + i);
            if constexpr (PRIMAL)
                EpiPrimalHal{stp, lam, 1.0f, A.c32, A.lx32, A.ux32, A.x0, A.x, A.xe, keep ? A.xt : nullptr}(r, s);
            else
                EpiDualHal{stp, lam, 1.0f, A.lr32, A.ur32, it.b.x ? A.g32 : nullptr, A.y0, A.y,
                           keep ? A.yt : nullptr}(r, s);
        }
    }
}

// Steps j0 .. j1-1 of NA slices at once (one row per lane in each): the
// 2 NA loads of a step are independent, so a wavefront keeps them in
// flight together, and U steps are unrolled (4 for the longer rows of A,
// 2 for A^T, measured).  Each row is summed in CSR order with FMA, as the
// one-lane-per-row kernels sum it.
template <int NA, int U>
__device__ __forceinline__ void slice_steps(const float* __restrict__ val, const uint16_t* __restrict__ idx,
                                            const float* xv, const int (&base)[4], int j0, int j1,
                                            float (&s)[4]) {
#pragma unroll U
    for (int j = j0; j < j1; ++j) {
        float v[NA];
        int c[NA];
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            v[u] = val[base[u] + j * kSliceRows];
            c[u] = idx[base[u] + j * kSliceRows];
        }
#pragma unroll
        for (int u = 0; u < NA; ++u) s[u] = fmaf(v[u], xv[c[u]], s[u]);
    }
}

// Sliced kernels: a work item is a group of up to kSliceGroup or
// kSliceGroupWide rows of A (kSliceGroupT of A^T, whose shorter rows gain
// from the larger group) of
// one LP in the sliced layout, walked in passes of 16 slices.  In a pass,
// wavefront w takes slices w, w + 4, w + 8 and w + 12 and walks them
// together: lane l reads entry j of its row in slice q at ptr[q] + j * 64
// + l, so each value and index load of 64 lanes is one contiguous 256-byte
// run.  Slice lengths do not grow
// along the group (its rows are sorted by length), so the walk runs all
// four slices up to the shortest one's length, then three, two, one.
// STAGED first copies the group's column window of the gathered vector (y
// for A^T, xe for A) into LDS with coalesced loads and gathers from there;
// otherwise the gathers go to global memory, up to 64 distinct cache lines
// per wavefront load.  The sums go to LDS at the row's place in the group,
// and the epilogue runs over the group's rows in their own order
// (coalesced vector accesses).
template <bool PRIMAL, bool STAGED, class VT>
__global__ __launch_bounds__(kBlock) void k_e_half(BatchArrays<VT> A, const WorkItem* __restrict__ work,
                                                   const int* __restrict__ wcnt, int i) {
    constexpr int kWaves = kBlock / 64, kPer = 4, kPass = kWaves * kPer;
    constexpr int kG = PRIMAL ? kSliceGroupT : kSliceGroupWide;  // the largest group of either side
    extern __shared__ float xs[];
    __shared__ float sums[kG];
    const int t = static_cast<int>(threadIdx.x);
    const int lane = t & 63, wv = t >> 6;
    const float* val = PRIMAL ? A.st_val : A.sa_val;
    const uint16_t* idx = PRIMAL ? A.st_idx : A.sa_idx;
    const int* ptr = PRIMAL ? A.st_ptr : A.sa_ptr;
    const int* len = PRIMAL ? A.st_len : A.sa_len;
    const int* pos = PRIMAL ? A.st_pos : A.sa_pos;
    const int2* win = PRIMAL ? A.st_win : A.sa_win;
    int w0, w1, step;
    xcd_slice(wcnt[PRIMAL ? 1 : 0], w0, w1, step);
    for (int w = w0; w < w1; w += step) {
        const ItemRegs it = load_item(work, w);
        const int row0 = it.a.x, cnt = it.a.y, chunk = it.a.w, first = it.b.w;
        if (i >= chunk) continue;  // whole item: uniform over the workgroup
        const int nsl = (cnt + kSliceRows - 1) / kSliceRows;
        const int2 wn = win[first];
        const float* xv = (PRIMAL ? A.y : A.xe) + it.b.y + wn.x;
        if constexpr (STAGED) {
#pragma unroll 4
            for (int q = t; q < wn.y; q += kBlock) xs[q] = xv[q];
            __syncthreads();
            xv = xs;
        }
        for (int p0 = 0; p0 < nsl; p0 += kPass) {
            int base[kPer], ln[kPer], ps[kPer];
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int q = p0 + wv + kWaves * u;
                const bool on = q < nsl;
                base[u] = on ? ptr[first + q] + lane : 0;
                ln[u] = on ? len[first + q] : 0;
                ps[u] = on ? pos[(first + q) * kSliceRows + lane] : -1;
            }
            float sm[kPer] = {0.0f, 0.0f, 0.0f, 0.0f};
            slice_steps<4, PRIMAL ? 2 : 4>(val, idx, xv, base, 0, ln[3], sm);
            slice_steps<3, PRIMAL ? 2 : 4>(val, idx, xv, base, ln[3], ln[2], sm);
            slice_steps<2, PRIMAL ? 2 : 4>(val, idx, xv, base, ln[2], ln[1], sm);
            slice_steps<1, PRIMAL ? 2 : 4>(val, idx, xv, base, ln[1], ln[0], sm);
#pragma unroll
            for (int u = 0; u < kPer; ++u)
                if (ps[u] >= 0) sums[ps[u]] = sm[u];
        }
        __syncthreads();
        const bool keep = i == chunk - 1;
        const float stp = __int_as_float(it.b.z), lam = Control::lambda(it.a.z + i);
        for (int r = t; r < cnt; r += kBlock) {
            if constexpr (PRIMAL)
                EpiPrimalHal{stp, lam, 1.0f, A.c32, A.lx32, A.ux32, A.x0, A.x, A.xe, keep ? A.xt : nullptr}(
                    row0 + r, sums[r]);
            else
                EpiDualHal{stp, lam, 1.0f, A.lr32, A.ur32, it.b.x ? A.g32 : nullptr, A.y0, A.y,
                           keep ? A.yt : nullptr}(row0 + r, sums[r]);
        }
        __syncthreads();
    }
}

// ---- round and check phases over the work items --------------------------
//
// The FP64 products of a refinement round's data and of the check's KKT
// error run over blocks of kCheckRows rows of every LP with a chunk in
// flight, so an LP's products are spread over the chip however few LPs
// still run, and the check's sums do not depend on the iterate kernels.
// Rows take lanes_for(mean row length) lanes each.

enum ItemPass { kPassRound, kPassTotals, kPassKkt };

// f(row in the item, lane, lanes as an integral constant) for the item's
// rows, L lanes per row.
template <int L, class F>
__device__ __forceinline__ void item_rows(int cnt, const F& f) {
    constexpr int G = kBlock / L;
    const int g = static_cast<int>(threadIdx.x) / L, lane = static_cast<int>(threadIdx.x) & (L - 1);
    for (int r0 = 0; r0 < cnt; r0 += G)
        if (r0 + g < cnt) f(r0 + g, lane, std::integral_constant<int, L>{});
}

template <class F>
__device__ __forceinline__ void item_rows_any(int L, int cnt, const F& f) {
    switch (L) {
        case 1: item_rows<1>(cnt, f); break;
        case 2: item_rows<2>(cnt, f); break;
        case 4: item_rows<4>(cnt, f); break;
        case 8: item_rows<8>(cnt, f); break;
        case 16: item_rows<16>(cnt, f); break;
        default: item_rows<32>(cnt, f); break;
    }
}

// One pass over the blocks of both lists (A rows first, then A^T rows):
//  kPassRound:  for LPs starting a refinement round, the correction-LP data
//               from FP64 A x_base and A^T y_base, iterates reset;
//  kPassTotals: the total iterates x_tot, y_tot of the chunk's last
//               iterate (base point plus the scaled correction);
//  kPassKkt:    FP64 A x_tot and A^T y_tot with the relative-KKT terms,
//               reduced per block into part_r / part_c (kPart doubles each).
template <int PASS, class VT>
__global__ __launch_bounds__(kBlock) void k_b_items(BatchArrays<VT> A, const BatchState* __restrict__ S) {
    __shared__ double red[16];
    const int tid = static_cast<int>(threadIdx.x);
    const int nr = A.wcnt[2], total = nr + A.wcnt[3];
    int w0, w1, step;
    xcd_slice(total, w0, w1, step);
    for (int w = w0; w < w1; w += step) {
        const bool rows = w < nr;
        const int wi = rows ? w : w - nr;
        const int2 ck = rows ? A.ck_r[wi] : A.ck_c[wi];
        const int lp = ck.x;
        const BatchDesc& d = A.desc[lp];
        const View<VT> v = view_of(A, d);
        const int first = ck.y * kCheckRows;
        const int cnt = min(kCheckRows, (rows ? d.m : d.n) - first);
        if constexpr (PASS == kPassRound) {
            if (!S[lp].round_start) continue;  // uniform over the workgroup
            const double dp = S[lp].ctl.DP, dd = S[lp].ctl.DD;
            const int ph = S[lp].pol.phase;
            const int mode = ph == Polish::PRIMAL ? kRoundPrimalFeas : (ph == Polish::DUAL ? kRoundDualFeas : kRoundLp);
            if (rows) {
                const EpiRoundRow epi{v.lr, v.ur, v.ybase, v.dr, dp, dd, v.lr32, v.ur32, v.g32, mode};
                item_rows_any(v.la, cnt, [&](int r, int lane, auto LC) {
                    const int i = first + r;
                    const double ax =
                        row_dot<decltype(LC)::value, false, double>(v.rp[i], v.rp[i + 1], lane, v.ci, v.oa, v.xbase);
                    if (lane == 0) {
                        epi(i, ax);
                        v.y[i] = v.y0[i] = v.yl[i] = 0.0f;
                    }
                });
            } else {
                const EpiRoundCol epi{v.cmin, v.lx, v.ux, v.xbase, v.dc, dp, dd, v.c32, v.lx32, v.ux32, mode};
                item_rows_any(v.lat, cnt, [&](int r, int lane, auto LC) {
                    const int j = first + r;
                    const double aty =
                        row_dot<decltype(LC)::value, false, double>(v.cp[j], v.cp[j + 1], lane, v.ri, v.oat, v.ybase);
                    if (lane == 0) {
                        epi(j, aty);
                        const float x0v = fmaxf(v.lx32[j], fminf(0.0f, v.ux32[j]));
                        v.x[j] = v.x0[j] = v.xl[j] = x0v;
                    }
                });
            }
        } else if constexpr (PASS == kPassTotals) {
            if (S[lp].pol.phase == Polish::VERIFY) continue;  // x_tot, y_tot hold the polished pair
            if (rows) {
                const double invd = 1.0 / S[lp].ctl.DD;
                for (int r = tid; r < cnt; r += kBlock) {
                    const int i = first + r;
                    double t = v.ybase[i] + v.dr[i] * static_cast<double>(v.yt[i]) * invd;
                    if (!(v.ur[i] < kInf) && t > 0.0) t = 0.0;
                    if (!(v.lr[i] > -kInf) && t < 0.0) t = 0.0;
                    v.ytot[i] = t;
                }
            } else {
                const double invp = 1.0 / S[lp].ctl.DP;
                for (int r = tid; r < cnt; r += kBlock) {
                    const int j = first + r;
                    double t = v.xbase[j] + v.dc[j] * static_cast<double>(v.xt[j]) * invp;
                    t = t < v.lx[j] ? v.lx[j] : (t > v.ux[j] ? v.ux[j] : t);
                    v.xtot[j] = t;
                }
            }
        } else {
            if (rows) {
                const EpiKktRow epi{v.lr, v.ur, v.ytot, v.dr};
                double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                item_rows_any(v.la, cnt, [&](int r, int lane, auto LC) {
                    const int i = first + r;
                    const double ax =
                        row_dot<decltype(LC)::value, false, double>(v.rp[i], v.rp[i + 1], lane, v.ci, v.oa, v.xtot);
                    if (lane == 0) {
                        double t[7];
                        epi(i, ax, t);
#pragma unroll
                        for (int q = 0; q < 7; ++q) acc[q] = q < 6 ? acc[q] + t[q] : (t[q] > acc[q] ? t[q] : acc[q]);
                    }
                });
                block_reduce_all<6, 1, kBlock>(acc, red);
                if (tid < 7) A.part_r[static_cast<size_t>(wi) * kPart + tid] = red[tid];
            } else {
                const EpiKktCol epi{v.cmin, v.lx, v.ux, v.xtot, v.dc};
                double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                item_rows_any(v.lat, cnt, [&](int r, int lane, auto LC) {
                    const int j = first + r;
                    const double aty =
                        row_dot<decltype(LC)::value, false, double>(v.cp[j], v.cp[j + 1], lane, v.ri, v.oat, v.ytot);
                    if (lane == 0) {
                        double t[6];
                        epi(j, aty, t);
#pragma unroll
                        for (int q = 0; q < 6; ++q) acc[q] = q < 5 ? acc[q] + t[q] : (t[q] > acc[q] ? t[q] : acc[q]);
                    }
                });
                block_reduce_all<5, 1, kBlock>(acc, red);
                if (tid < 6) A.part_c[static_cast<size_t>(wi) * kPart + tid] = red[tid];
            }
        }
    }
}

// What a check does to the arrays of an LP besides the Control actions:
// the steps of a feasibility polish (Polish in control.hpp).
enum PolishEvent {
    kEvNone = 0,
    kEvStartPrimal,   // save the candidate; primal phase from x_cand
    kEvStartDual,     // save the candidate (x_cand is already polished); dual phase from y_cand
    kEvPrimalToDual,  // keep the polished x; dual phase from the saved y
    kEvCombine,       // polished x with the saved y (primal phase done, dual not needed): verify next
    kEvCombineDual,   // polished x with the polished y (dual phase done): verify next
    kEvResume,        // back to the main solve, restored as it was when the polish began
    kEvPhaseRestart,  // restart inside a phase
    kEvLimit          // budget spent inside a phase: the saved candidate is the result
};

// Phase 3, per LP: the KKT error from its blocks' partial sums -> Control
// decision -> restart / refine / polish / finish.  Counts the LPs still
// active.
template <class VT>
__global__ __launch_bounds__(kBB) void k_b_check(BatchArrays<VT> A, BatchState* S, int* active_count) {
    const int lp = A.order[blockIdx.x];
    BatchState& st = S[lp];
    if (!st.active) return;
    BatchDesc& d = A.desc[lp];
    const View<VT> v = view_of(A, d);
    const int m = v.m, n = v.n, tid = threadIdx.x;
    __shared__ double red[16], keep[8];
    __shared__ int act_s, ev_s;
    __shared__ Control ctl;
    __shared__ Polish pol;
    if (tid == 0) {
        ctl = st.ctl;
        pol = st.pol;
    }
    __syncthreads();
    const int chunk = A.it[lp].chunk;
    if (tid == 0) st.round_start = 0;  // the round pass of this chunk has run
    int act, ev = kEvNone;
    if (chunk > 0) {
        // The KKT terms of x_tot, y_tot, reduced per block by the check
        // pass, summed over the LP's blocks in order.
        const int4 wr = A.wrange[lp];
        if (tid < 7) {
            double a = 0.0;
            for (int q = 0; q < wr.y; ++q) {
                const double t = A.part_r[static_cast<size_t>(wr.x + q) * kPart + tid];
                a = tid < 6 ? a + t : (t > a ? t : a);
            }
            keep[tid] = a;
        } else if (tid >= 64 && tid < 70) {
            const int j = tid - 64;
            double a = 0.0;
            for (int q = 0; q < wr.w; ++q) {
                const double t = A.part_c[static_cast<size_t>(wr.z + q) * kPart + j];
                a = j < 5 ? a + t : (t > a ? t : a);
            }
            red[j] = a;
        }
        __syncthreads();
        if (tid == 0) {
            const double rel = kkt_finish(keep, red, d);
            ctl.advance(chunk);
            int a = Control::CONTINUE, e = kEvNone;
            if (pol.phase == Polish::MAIN) {
                a = ctl.on_check(rel, d.rel_p, d.rel_d, d.rel_g);
                st.have_cand = 1;
                if (a == Control::CONTINUE && pol.trigger(A.tol, ctl.total_it, rel, d.rel_g)) {
                    st.ctl_held = ctl;
                    pol.save(ctl.total_it, rel, d.rel_p, d.rel_d);
                    const bool primal = pol.needs(A.tol, d.rel_p);
                    pol.start(primal ? Polish::PRIMAL : Polish::DUAL, ctl.total_it);
                    ctl.DP = ctl.DD = Polish::scale(primal ? d.rel_p : d.rel_d, ctl.DP);
                    if (!primal) ctl.use_g = true;
                    e = primal ? kEvStartPrimal : kEvStartDual;
                }
            } else if (pol.phase == Polish::VERIFY) {
                // the polished pair's KKT terms came from the grid-wide passes
                pol.phase = Polish::MAIN;
                if (Polish::accept(A.tol, rel, d.rel_g)) a = Control::OPTIMAL;
                else e = ctl.total_it >= ctl.max_iterations ? kEvLimit : kEvResume;
            } else {
                const bool dual_phase = pol.phase == Polish::DUAL;
                const int pa = pol.on_check(A.tol, dual_phase ? d.rel_d : d.rel_p, chunk);
                const bool spent = ctl.total_it >= ctl.max_iterations;
                if (pa == Polish::DONE && (dual_phase || !pol.needs(A.tol, pol.save_d)) && !spent) {
                    e = dual_phase ? kEvCombineDual : kEvCombine;
                    pol.phase = Polish::VERIFY;
                } else if (spent) {
                    e = kEvLimit;
                } else if (pa == Polish::DONE) {
                    pol.start(Polish::DUAL, ctl.total_it);
                    ctl.DP = ctl.DD = Polish::scale(pol.save_d, ctl.DP);
                    ctl.use_g = true;
                    e = kEvPrimalToDual;
                } else if (pa == Polish::FAILED) {
                    e = kEvResume;
                } else if (pa == Polish::RESTART) {
                    e = kEvPhaseRestart;
                }
            }
            act_s = a;
            ev_s = e;
        }
        __syncthreads();
        act = act_s;
        ev = ev_s;
    } else {
        act = Control::LIMIT;
        if (pol.phase != Polish::MAIN) ev = kEvLimit;
    }
    // Polish steps.
    if (ev == kEvStartPrimal || ev == kEvStartDual) {
        // hold the main solve's state (iterates, round data, round base)
        float* const mx[6] = {v.x, v.x0, v.xl, v.c32, v.lx32, v.ux32};
        float* const my[6] = {v.y, v.y0, v.yl, v.lr32, v.ur32, v.g32};
        for (int j = tid; j < n; j += kBB) {
#pragma unroll
            for (int q = 0; q < 6; ++q) v.hx[q][j] = mx[q][j];
            v.hxb[j] = v.xbase[j];
        }
        for (int i = tid; i < m; i += kBB) {
#pragma unroll
            for (int q = 0; q < 6; ++q) v.hy[q][i] = my[q][i];
            v.hyb[i] = v.ybase[i];
        }
        __syncthreads();
        for (int j = tid; j < n; j += kBB) {
            const double x = v.xtot[j];
            v.xsave[j] = x;
            v.xbase[j] = ev == kEvStartPrimal ? x : 0.0;
            if (ev == kEvStartDual) v.xpol[j] = x;
        }
        for (int i = tid; i < m; i += kBB) {
            const double y = v.ytot[i];
            v.ysave[i] = y;
            v.ybase[i] = ev == kEvStartPrimal ? 0.0 : y;
        }
        if (tid == 0) st.round_start = 1;
    } else if (ev == kEvPrimalToDual) {
        for (int j = tid; j < n; j += kBB) {
            v.xpol[j] = v.xtot[j];
            v.xbase[j] = 0.0;
        }
        for (int i = tid; i < m; i += kBB) v.ybase[i] = v.ysave[i];
        if (tid == 0) st.round_start = 1;
    } else if (ev == kEvCombine || ev == kEvCombineDual) {
        // (x_tot is the polished x after a primal phase; y_tot the polished y after a dual one)
        if (ev == kEvCombine)
            for (int i = tid; i < m; i += kBB) v.ytot[i] = v.ysave[i];
        else
            for (int j = tid; j < n; j += kBB) v.xtot[j] = v.xpol[j];
    } else if (ev == kEvPhaseRestart) {
        for (int j = tid; j < n; j += kBB) v.x[j] = v.x0[j] = v.xl[j] = v.xt[j];
        for (int i = tid; i < m; i += kBB) v.y[i] = v.y0[i] = v.yl[i] = v.yt[i];
    }
    if (ev == kEvResume || ev == kEvLimit) {
        for (int j = tid; j < n; j += kBB) v.xtot[j] = v.xsave[j];
        for (int i = tid; i < m; i += kBB) v.ytot[i] = v.ysave[i];
        if (ev == kEvResume) {
            // the main solve continues where it was: the polish cost only its iterations
            float* const mx[6] = {v.x, v.x0, v.xl, v.c32, v.lx32, v.ux32};
            float* const my[6] = {v.y, v.y0, v.yl, v.lr32, v.ur32, v.g32};
            for (int j = tid; j < n; j += kBB) {
#pragma unroll
                for (int q = 0; q < 6; ++q) mx[q][j] = v.hx[q][j];
                v.xbase[j] = v.hxb[j];
            }
            for (int i = tid; i < m; i += kBB) {
#pragma unroll
                for (int q = 0; q < 6; ++q) my[q][i] = v.hy[q][i];
                v.ybase[i] = v.hyb[i];
            }
        }
        __syncthreads();
        if (ev == kEvLimit) {
            (void)batch_kkt(v, d, red, keep);  // the result fields describe the saved candidate
            act = Control::LIMIT;
        } else if (tid == 0) {
            const int spent = ctl.total_it;
            ctl = st.ctl_held;
            ctl.total_it = spent;
            st.have_cand = 1;
        }
        if (tid == 0) {
            pol.phase = Polish::MAIN;
            ++pol.fails;
        }
    }
    bool done = act == Control::OPTIMAL || act == Control::FAIL || act == Control::LIMIT;
    if (act == Control::REFINE) {
        for (int j = tid; j < n; j += kBB) v.xbase[j] = v.xtot[j];
        for (int i = tid; i < m; i += kBB) v.ybase[i] = v.ytot[i];
        if (tid == 0) {
            st.round_start = 1;
            st.have_cand = 0;
        }
    } else if (act == Control::RESTART) {
        double acc[2] = {0.0, 0.0};
        for (int j = tid; j < n; j += kBB) {
            const double df = static_cast<double>(v.xt[j]) - v.xl[j];
            acc[0] += df * df;
            v.x[j] = v.x0[j] = v.xl[j] = v.xt[j];
        }
        for (int i = tid; i < m; i += kBB) {
            const double df = static_cast<double>(v.yt[i]) - v.yl[i];
            acc[1] += df * df;
            v.y[i] = v.y0[i] = v.yl[i] = v.yt[i];
        }
        block_reduce_all<2, 0>(acc, red);
        if (tid == 0) ctl.on_restart(sqrt(red[0]), sqrt(red[1]));
    }
    if (done) {
        __syncthreads();
        if (!st.have_cand) {  // the budget ran out right after a refinement
            for (int j = tid; j < n; j += kBB) v.xtot[j] = v.xbase[j];
            for (int i = tid; i < m; i += kBB) v.ytot[i] = v.ybase[i];
            __syncthreads();
            const double rel = batch_kkt(v, d, red, keep);
            if (tid == 0 && rel <= A.tol) act = Control::OPTIMAL;
        }
        if (tid == 0) {
            d.status = act == Control::OPTIMAL ? PM_SOLVE_OPTIMAL
                       : act == Control::FAIL  ? PM_SOLVE_NUMERICAL_FAILURE
                                               : PM_SOLVE_ITERATION_LIMIT;
            d.iterations = ctl.total_it;
            d.restarts = ctl.restarts;
            d.refinements = ctl.rounds;
            d.polishes = pol.attempts;
            d.omega = ctl.omega;
            d.t_end = wall_clock64();
            st.active = 0;
        }
    }
    if (tid == 0) {
        st.ctl = ctl;
        st.pol = pol;
        publish_iter(A.it, lp, st, d, A.check_frequency);
        if (!done) atomicAdd(active_count, 1);
    }
}

__global__ void k_zero_int(int* p) { *p = 0; }

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(int64_t n, const float* __restrict__ a,
                                                       double* __restrict__ b) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) b[i] = static_cast<double>(a[i]);
}

// out[e] = v[src[e]], 0 where src[e] < 0.
__global__ __launch_bounds__(kBlock) void k_take_f32(int64_t n, const int* __restrict__ src,
                                                     const float* __restrict__ v, float* __restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (e < n) out[e] = src[e] >= 0 ? v[src[e]] : 0.0f;
}

}  // namespace
}  // namespace pm

using namespace pm;

namespace pm {
namespace {

void batch_autotune(pm_batch* b);
void set_mode(pm_batch* b, bool tail);

void batch_create(pm_batch* b, int count, const pm_lp* lps, const pm_pdhg_options* o) {
    if (count <= 0) fail(PM_ERR_INVALID_ARG, "batch count must be positive");
    if (!lps) fail(PM_ERR_INVALID_ARG, "lps is NULL");
    pm_pdhg_options opts;
    pm_pdhg_options_default(&opts);
    if (o) opts = *o;
    if (opts.check_frequency <= 0) fail(PM_ERR_INVALID_ARG, "check_frequency must be positive");
    if (opts.step_rule != PM_STEP_HALPERN)
        fail(PM_ERR_INVALID_ARG, "the batch engine runs the reflected Halpern iteration (PM_STEP_HALPERN)");
    if (opts.restart_metric != PM_RESTART_KKT)
        fail(PM_ERR_INVALID_ARG, "the batch engine restarts on the relative KKT error (PM_RESTART_KKT)");
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device) == hipSuccess && ncu > 0)
        b->num_cu = ncu;
    b->tol = opts.tolerance;
    b->max_iterations = opts.max_iterations;
    b->check_frequency = opts.check_frequency;
    b->max_refinements = opts.max_refinements;
    b->polish = opts.feasibility_polishing ? 1 : 0;
    b->count = count;
    b->desc.assign(count, BatchDesc{});
    b->sense.resize(count);
    b->offset.resize(count);
    // Offsets and validation.
    int64_t rpo = 0, cpo = 0, nzo = 0, ro = 0, co = 0;
    for (int k = 0; k < count; ++k) {
        const pm_lp& L = lps[k];
        if (L.m < 0 || L.n <= 0 || L.nnz < 0) fail(PM_ERR_INVALID_ARG, "batch LP " + std::to_string(k) + ": bad size");
        if (!L.row_ptr || (L.nnz > 0 && (!L.col_idx || !L.values)) || !L.c || !L.var_lower || !L.var_upper ||
            (L.m > 0 && (!L.row_lower || !L.row_upper)))
            fail(PM_ERR_INVALID_ARG, "batch LP " + std::to_string(k) + ": NULL array");
        if (L.row_ptr[0] != 0 || L.row_ptr[L.m] != L.nnz)
            fail(PM_ERR_INVALID_ARG, "batch LP " + std::to_string(k) + ": row_ptr inconsistent with nnz");
        for (int i = 0; i < L.m; ++i)
            if (L.row_ptr[i] > L.row_ptr[i + 1]) fail(PM_ERR_INVALID_ARG, "batch LP: row_ptr decreasing");
        for (int q = 0; q < L.nnz; ++q)
            if (L.col_idx[q] < 0 || L.col_idx[q] >= L.n) fail(PM_ERR_INVALID_ARG, "batch LP: col_idx out of range");
        for (int j = 0; j < L.n; ++j)
            if (!std::isfinite(L.c[j])) fail(PM_ERR_INVALID_ARG, "batch LP: non-finite cost");
        BatchDesc& d = b->desc[k];
        d.m = L.m;
        d.n = L.n;
        d.nnz = L.nnz;
        d.rp_off = rpo;
        d.cp_off = cpo;
        d.nz_off = nzo;
        d.row_off = ro;
        d.col_off = co;
        rpo += L.m + 1;
        cpo += L.n + 1;
        nzo += L.nnz;
        ro += L.m;
        co += L.n;
        b->sense[k] = L.sense;
        b->offset[k] = L.offset;
        for (int q = 0; q < L.nnz && b->exact; ++q)
            if (static_cast<double>(static_cast<float>(L.values[q])) != L.values[q]) b->exact = false;
    }
    if (nzo >= INT32_MAX || ro >= INT32_MAX || co >= INT32_MAX)
        fail(PM_ERR_INVALID_ARG, "batch too large: rows, columns and nonzeros must each stay below 2^31");
    b->rows_total = ro;
    b->cols_total = co;
    b->nnz_total = nzo;
    const size_t R = std::max<int64_t>(ro, 1), Cn = std::max<int64_t>(co, 1), Z = std::max<int64_t>(nzo, 1);
    std::vector<int> h_rp(rpo), h_cp(cpo), h_ci(Z), h_ri(Z);
    std::vector<int> h_gci(Z), h_gri(Z);  // global column / row index of each entry
    std::vector<float> h_oa32, h_oat32;
    std::vector<double> h_oa64, h_oat64;
    if (b->exact) { h_oa32.resize(Z); h_oat32.resize(Z); }
    else { h_oa64.resize(Z); h_oat64.resize(Z); }
    std::vector<double> h_cmin(Cn), h_lx(Cn), h_ux(Cn), h_lr(R), h_ur(R);
    std::vector<int> h_grp(R + 1, 0), h_gcp(Cn + 1, 0);
    std::vector<std::vector<int2>> seg_a_lp(count), seg_at_lp(count);
    std::vector<SlicedHost> sl_a_lp(count), sl_at_lp(count);
    h_grp[ro] = static_cast<int>(nzo);
    h_gcp[co] = static_cast<int>(nzo);
    b->c_model.assign(Cn, 0.0);
    // Per-LP preparation (transpose, layouts) on host threads; the scaling
    // runs on the device afterwards.
    auto prep = [&](int k) {
        const pm_lp& L = lps[k];
        BatchDesc& d = b->desc[k];
        const int m = L.m, n = L.n, nnz = L.nnz;
        std::vector<int> cp, ri;
        std::vector<int64_t> perm;
        csr_transpose(m, n, L.row_ptr, L.col_idx, cp, ri, perm);
        std::copy(L.row_ptr, L.row_ptr + m + 1, h_rp.begin() + d.rp_off);
        std::copy(cp.begin(), cp.end(), h_cp.begin() + d.cp_off);
        for (int i = 0; i < m; ++i) {
            h_grp[d.row_off + i] = static_cast<int>(d.nz_off + L.row_ptr[i]);
        }
        for (int j = 0; j < n; ++j) {
            h_gcp[d.col_off + j] = static_cast<int>(d.nz_off + cp[j]);
        }
        for (int q = 0; q < nnz; ++q) {
            h_ci[d.nz_off + q] = L.col_idx[q];
            h_ri[d.nz_off + q] = ri[q];
            h_gci[d.nz_off + q] = static_cast<int>(d.col_off + L.col_idx[q]);
            h_gri[d.nz_off + q] = static_cast<int>(d.row_off + ri[q]);
            if (b->exact) {
                h_oa32[d.nz_off + q] = static_cast<float>(L.values[q]);
                h_oat32[d.nz_off + q] = static_cast<float>(L.values[perm[q]]);
            } else {
                h_oa64[d.nz_off + q] = L.values[q];
                h_oat64[d.nz_off + q] = L.values[perm[q]];
            }
        }
        const double sg = L.sense == 1 ? -1.0 : 1.0;
        for (int j = 0; j < n; ++j) {
            h_cmin[d.col_off + j] = sg * L.c[j];
            b->c_model[d.col_off + j] = L.c[j];
            h_lx[d.col_off + j] = L.var_lower[j];
            h_ux[d.col_off + j] = L.var_upper[j];
        }
        for (int i = 0; i < m; ++i) {
            h_lr[d.row_off + i] = L.row_lower[i];
            h_ur[d.row_off + i] = L.row_upper[i];
        }
        d.lanes_a = m > 0 ? lanes_for(static_cast<double>(nnz) / m) : 1;
        d.lanes_at = lanes_for(static_cast<double>(nnz) / n);
        stream_segments(m, L.row_ptr, seg_a_lp[k]);
        stream_segments(n, cp.data(), seg_at_lp[k]);
        d.ga = balanced_group(m, group_rows_a(m, n, L.row_ptr, L.col_idx));
        d.bucket = size_class(m);
        d.gt = balanced_group(n, kSliceGroupT);
        build_sliced(m, L.row_ptr, L.col_idx, d.nz_off, sl_a_lp[k], d.ga);
        build_sliced(n, cp.data(), ri.data(), d.nz_off, sl_at_lp[k], d.gt);
    };
    {
        const int nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        std::atomic<int> next{0};
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&] {
                for (int k = next++; k < count; k = next++) prep(k);
            });
        for (auto& t : th) t.join();
    }
    // Stream segments, concatenated in LP order.
    {
        std::vector<int2> sa, st;
        for (int k = 0; k < count; ++k) {
            BatchDesc& d = b->desc[k];
            d.sa_off = static_cast<int32_t>(sa.size());
            d.sa_cnt = static_cast<int32_t>(seg_a_lp[k].size());
            d.st_off = static_cast<int32_t>(st.size());
            d.st_cnt = static_cast<int32_t>(seg_at_lp[k].size());
            sa.insert(sa.end(), seg_a_lp[k].begin(), seg_a_lp[k].end());
            st.insert(st.end(), seg_at_lp[k].begin(), seg_at_lp[k].end());
        }
        if (sa.empty()) sa.push_back(make_int2(0, 0));
        if (st.empty()) st.push_back(make_int2(0, 0));
        b->seg_a.upload(sa, b->stream);
        b->seg_at.upload(st, b->stream);
    }
    // Sliced copies, concatenated in LP order with global entry offsets.
    DevBuf<int> src_a, src_at;  // slot -> CSR entry, until the values are filled
    auto concat = [&](std::vector<SlicedHost>& parts, bool a_side, SlicedCsr& dst, DevBuf<int>& src) {
        SlicedHost all;
        for (int k = 0; k < count; ++k) {
            SlicedHost& p = parts[k];
            (a_side ? b->desc[k].la_off : b->desc[k].lt_off) = static_cast<int32_t>(all.len.size());
            const int64_t shift = static_cast<int64_t>(all.src.size());
            for (int64_t q : p.ptr) all.ptr.push_back(q + shift);
            all.len.insert(all.len.end(), p.len.begin(), p.len.end());
            all.pos.insert(all.pos.end(), p.pos.begin(), p.pos.end());
            all.win.insert(all.win.end(), p.win.begin(), p.win.end());
            all.src.insert(all.src.end(), p.src.begin(), p.src.end());
            all.idx.insert(all.idx.end(), p.idx.begin(), p.idx.end());
            SlicedHost().src.swap(p.src);
            SlicedHost().idx.swap(p.idx);
        }
        if (all.src.size() >= static_cast<size_t>(INT32_MAX))
            fail(PM_ERR_INVALID_ARG, "batch too large for the sliced layout");
        std::vector<int> ptr32(all.ptr.begin(), all.ptr.end());
        int& widest = a_side ? b->win_a : b->win_at;
        for (const int2& w : all.win) widest = std::max(widest, w.y);
        if (ptr32.empty()) { ptr32.push_back(0); all.len.push_back(0); all.win.push_back(make_int2(0, 0)); }
        if (all.pos.empty()) all.pos.assign(kSliceRows, -1);
        if (all.src.empty()) { all.src.push_back(-1); all.idx.push_back(0); }
        src.upload(all.src, b->stream);
        dst.val.resize(all.src.size());
        dst.idx.upload(all.idx, b->stream);
        dst.ptr.upload(ptr32, b->stream);
        dst.len.upload(all.len, b->stream);
        dst.pos.upload(all.pos, b->stream);
        dst.win.upload(all.win, b->stream);
        PM_HIP(hipStreamSynchronize(b->stream));  // the host vectors die here
    };
    concat(sl_a_lp, true, b->sl_a, src_a);
    concat(sl_at_lp, false, b->sl_at, src_at);
    // Largest LP first.
    b->order.resize(count);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
        return b->desc[a].nnz + b->desc[a].m + b->desc[a].n > b->desc[c].nnz + b->desc[c].m + b->desc[c].n;
    });
    hipStream_t s = b->stream;
    b->rp.upload(h_rp, s);
    b->cp.upload(h_cp, s);
    b->ci.upload(h_ci, s);
    b->ri.upload(h_ri, s);
    if (b->exact) { b->oa32.upload(h_oa32, s); b->oat32.upload(h_oat32, s); }
    else { b->oa64.upload(h_oa64, s); b->oat64.upload(h_oat64, s); }
    b->cmin.upload(h_cmin, s);
    b->lx.upload(h_lx, s);
    b->ux.upload(h_ux, s);
    b->lr.upload(h_lr, s);
    b->ur.upload(h_ur, s);
    b->d_order.upload(b->order, s);
    b->grp.upload(h_grp, s);
    b->gcp.upload(h_gcp, s);
    // Ruiz + Pock-Chambolle scaling of the block-diagonal concatenation on
    // the device (each LP's factors are those of the LP alone), the FP32
    // iterate values, and the values of the sliced layouts.
    {
        DevBuf<double> vs(Z), vst(Z), rn(R), cn(Cn);
        DevBuf<int> gci, gri;
        gci.upload(h_gci, s);
        gri.upload(h_gri, s);
        const dim3 gz = grid_for(static_cast<int64_t>(Z)), blk(kBlock);
        if (b->exact) {
            hipLaunchKernelGGL(k_f32_to_f64, gz, blk, 0, s, static_cast<int64_t>(Z), b->oa32.get(), vs.get());
            hipLaunchKernelGGL(k_f32_to_f64, gz, blk, 0, s, static_cast<int64_t>(Z), b->oat32.get(), vst.get());
        } else {
            PM_HIP(hipMemcpyAsync(vs.get(), b->oa64.get(), sizeof(double) * Z, hipMemcpyDeviceToDevice, s));
            PM_HIP(hipMemcpyAsync(vst.get(), b->oat64.get(), sizeof(double) * Z, hipMemcpyDeviceToDevice, s));
        }
        b->dr.resize(R);
        b->dc.resize(Cn);
        ruiz_pc_device(s, opts.ruiz_iterations, static_cast<int>(ro), static_cast<int>(co), Z, b->grp.get(),
                       gci.get(), b->gcp.get(), gri.get(), vs.get(), vst.get(), b->dr.get(), b->dc.get(), rn.get(),
                       cn.get());
        b->va.resize(Z);
        b->vat.resize(Z);
        hipLaunchKernelGGL(k_to_f32, gz, blk, 0, s, static_cast<int64_t>(Z), vs.get(), b->va.get());
        hipLaunchKernelGGL(k_to_f32, gz, blk, 0, s, static_cast<int64_t>(Z), vst.get(), b->vat.get());
        const int64_t na = static_cast<int64_t>(b->sl_a.val.size()), nat = static_cast<int64_t>(b->sl_at.val.size());
        hipLaunchKernelGGL(k_take_f32, grid_for(na), blk, 0, s, na, src_a.get(), b->va.get(), b->sl_a.val.get());
        hipLaunchKernelGGL(k_take_f32, grid_for(nat), blk, 0, s, nat, src_at.get(), b->vat.get(), b->sl_at.val.get());
        PM_LAUNCHED();
        PM_HIP(hipStreamSynchronize(s));
    }
    b->iter.resize(static_cast<size_t>(count));
    // One lane count per product for the whole concatenation (its mean row
    // length), four chunks per graph unless that makes the graph large.
    b->lanes_a = ro > 0 ? lanes_for(static_cast<double>(nzo) / static_cast<double>(ro)) : 1;
    b->lanes_at = lanes_for(static_cast<double>(nzo) / static_cast<double>(co));
    // Work-list capacity: the largest lane count (the smallest groups) or
    // the stream segments, whichever gives more items.
    // Each LP's items under any kernel fit max(its groups of the largest
    // lane count, its stream segments); the size classes take consecutive
    // regions, and the list over every LP uses the whole array.
    constexpr int kMinGroup = kBlock / kMaxLanes;
    int64_t cap_r[pm_batch::kBuckets] = {0, 0, 0}, cap_c[pm_batch::kBuckets] = {0, 0, 0};
    for (const BatchDesc& d : b->desc) {
        cap_r[d.bucket] += std::max<int64_t>((d.m + kMinGroup - 1) / kMinGroup, d.sa_cnt);
        cap_c[d.bucket] += std::max<int64_t>((d.n + kMinGroup - 1) / kMinGroup, d.st_cnt);
        ++b->bucket_lps[d.bucket];
    }
    int64_t tot_r = 0, tot_c = 0;
    for (int c = 0; c < pm_batch::kBuckets; ++c) {
        b->off_r[c] = tot_r;
        b->off_c[c] = tot_c;
        tot_r += cap_r[c];
        tot_c += cap_c[c];
    }
    b->work_r.resize(static_cast<size_t>(std::max<int64_t>(tot_r, 1)));
    b->work_c.resize(static_cast<size_t>(std::max<int64_t>(tot_c, 1)));
    size_t blocks_r = 0, blocks_c = 0;
    for (const BatchDesc& d : b->desc) {
        blocks_r += (d.m + kCheckRows - 1) / kCheckRows;
        blocks_c += (d.n + kCheckRows - 1) / kCheckRows;
    }
    b->ck_r.resize(std::max<size_t>(blocks_r, 1));
    b->ck_c.resize(std::max<size_t>(blocks_c, 1));
    b->part_r.resize(b->ck_r.size() * kPart);
    b->part_c.resize(b->ck_c.size() * kPart);
    b->wrange.resize(static_cast<size_t>(count));
    b->wcnt.resize(4 + 2 * pm_batch::kBuckets);
    if (opts.spmv_variant == PM_SPMV_MERGE || opts.spmv_variant == PM_SPMV_COLBLOCK)
        fail(PM_ERR_INVALID_ARG, "spmv_variant PM_SPMV_MERGE and PM_SPMV_COLBLOCK are single-LP layouts");
    const int fixed = variant_kernel(opts.spmv_variant);
    if (fixed != kAutoKernel) b->lanes_a = b->lanes_at = fixed;
    for (int c = 0; c < pm_batch::kBuckets; ++c) {
        b->blanes_a[c] = b->lanes_a;
        b->blanes_at[c] = b->lanes_at;
    }
    if (fixed == kStagedKernel && std::max(b->win_a, b->win_at) > kStagedMaxVec)
        fail(PM_ERR_INVALID_ARG, "PM_SPMV_STAGED holds each group's column window in LDS: windows must stay within " +
                                     std::to_string(kStagedMaxVec) + " columns");
    if (fixed == kSlicedKernel && std::max(b->win_a, b->win_at) > kSlicedMaxWin)
        fail(PM_ERR_INVALID_ARG, "PM_SPMV_SLICED: a group's column window is wider than " +
                                     std::to_string(kSlicedMaxWin) + " columns");
    b->chunks_per_graph = std::max(1, std::min(4, 512 / std::max(1, b->check_frequency)));
    for (auto* p : {&b->c32, &b->lx32, &b->ux32, &b->x, &b->xe, &b->x0, &b->xt, &b->xl}) p->resize(Cn);
    for (auto* p : {&b->lr32, &b->ur32, &b->g32, &b->y, &b->y0, &b->yt, &b->yl}) p->resize(R);
    for (auto* p : {&b->xtot, &b->xbase, &b->wx, &b->xsave, &b->xpol}) p->resize(Cn);
    for (auto* p : {&b->ytot, &b->ybase, &b->wy, &b->ysave}) p->resize(R);
    b->hold32.resize(6 * (Cn + R));
    b->hold64.resize(Cn + R);
    b->d_desc.upload(b->desc, s);
    b->state.resize(sizeof(BatchState) * static_cast<size_t>(count));
    b->counter.resize(1);
    b->h_counter.resize(1);
    PM_HIP(hipStreamSynchronize(s));
    if (fixed == kAutoKernel) {
        b->tail_a = b->lanes_a;
        b->tail_at = b->lanes_at;
        b->tail_nnz = kTailNnz;
        batch_autotune(b);
    }
    set_mode(b, false);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess && khz > 0)
        b->clock_khz = khz;
}

template <class VT>
BatchArrays<VT> batch_arrays(pm_batch* b, const VT* oa, const VT* oat) {
    BatchArrays<VT> A{};
    A.rp = b->rp.get(); A.ci = b->ci.get(); A.va = b->va.get(); A.oa = oa;
    A.cp = b->cp.get(); A.ri = b->ri.get(); A.vat = b->vat.get(); A.oat = oat;
    A.cmin = b->cmin.get(); A.lr = b->lr.get(); A.ur = b->ur.get(); A.lx = b->lx.get(); A.ux = b->ux.get();
    A.dr = b->dr.get(); A.dc = b->dc.get(); A.wx = b->wx.get(); A.wy = b->wy.get();
    A.c32 = b->c32.get(); A.lr32 = b->lr32.get(); A.ur32 = b->ur32.get(); A.lx32 = b->lx32.get();
    A.ux32 = b->ux32.get(); A.g32 = b->g32.get();
    A.x = b->x.get(); A.xe = b->xe.get(); A.x0 = b->x0.get(); A.xt = b->xt.get(); A.xl = b->xl.get();
    A.y = b->y.get(); A.y0 = b->y0.get(); A.yt = b->yt.get(); A.yl = b->yl.get();
    A.xtot = b->xtot.get(); A.ytot = b->ytot.get(); A.xbase = b->xbase.get(); A.ybase = b->ybase.get();
    A.xsave = b->xsave.get(); A.ysave = b->ysave.get(); A.xpol = b->xpol.get();
    A.hold32 = b->hold32.get(); A.hold64 = b->hold64.get();
    A.cols_total = std::max<int64_t>(b->cols_total, 1); A.rows_total = std::max<int64_t>(b->rows_total, 1);
    A.desc = b->d_desc.get();
    A.order = b->d_order.get();
    A.grp = b->grp.get();
    A.gcp = b->gcp.get();
    A.it = b->iter.get();
    A.wrange = b->wrange.get(); A.ck_r = b->ck_r.get(); A.ck_c = b->ck_c.get();
    A.part_r = b->part_r.get(); A.part_c = b->part_c.get(); A.wcnt = b->wcnt.get();
    A.sa_val = b->sl_a.val.get(); A.sa_idx = b->sl_a.idx.get(); A.sa_ptr = b->sl_a.ptr.get();
    A.sa_len = b->sl_a.len.get(); A.sa_pos = b->sl_a.pos.get();
    A.st_val = b->sl_at.val.get(); A.st_idx = b->sl_at.idx.get(); A.st_ptr = b->sl_at.ptr.get();
    A.st_len = b->sl_at.len.get(); A.st_pos = b->sl_at.pos.get();
    A.sa_win = b->sl_a.win.get(); A.st_win = b->sl_at.win.get();
    A.tol = b->tol;
    A.max_iterations = b->max_iterations;
    A.check_frequency = b->check_frequency;
    A.max_refinements = b->max_refinements;
    A.polish = b->polish;
    return A;
}

// Workgroups (of kBlock threads, `lds` dynamic LDS bytes) of a persistent
// launch of `kernel`: as many as the device holds at once, a multiple of 8
// (xcd_slice).  Work items are assigned statically, so the workgroups of a
// larger grid that do not fit start only after others have run all of
// theirs: with items of even cost that costs time, with uneven ones it
// balances the load (the autotuner chooses, see mult_a / mult_at).
template <class K>
int resident_grid(pm_batch* b, K kernel, size_t lds) {
    const std::pair<const void*, size_t> key(reinterpret_cast<const void*>(kernel), lds);
    const auto it = b->grid_cache.find(key);
    if (it != b->grid_cache.end()) return it->second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, lds) != hipSuccess || per_cu <= 0)
        per_cu = b->iter_grid / std::max(1, b->num_cu);
    const int g = std::max(8, per_cu * b->num_cu / 8 * 8);
    b->grid_cache.emplace(key, g);
    return g;
}

template <class K, class... Args>
void launch_resident(pm_batch* b, int mult, K kernel, size_t lds, Args... args) {
    const int g = resident_grid(b, kernel, lds) * std::max(1, mult);
    hipLaunchKernelGGL(kernel, dim3(g), dim3(kBlock), lds, b->stream, args...);
}

// One iterate launch: half `which` (PM_KERNEL_SPMV_AT primal, PM_KERNEL_SPMV_A
// dual) of iteration i of the chunk, with the kernel for `lanes` (a lanes
// count or a kernel code), on `mult` resident grids.
template <class VT>
void launch_iterate(pm_batch* b, const BatchArrays<VT>& A, int which, int i, int lanes, int mult, int bucket = -1) {
    const bool primal = which == PM_KERNEL_SPMV_AT;
    if (!primal && b->rows_total == 0) return;
    WorkItem* work = primal ? b->work_c.get() : b->work_r.get();
    int* wc = b->wcnt.get();
    if (bucket >= 0) {
        work += primal ? b->off_c[bucket] : b->off_r[bucket];
        wc += 4 + 2 * bucket;
    }
    if (lanes == kStreamKernel) {
        if (primal) launch_resident(b, mult, k_t_half<true, VT>, 0, A, work, wc, i);
        else launch_resident(b, mult, k_t_half<false, VT>, 0, A, work, wc, i);
    } else if (lanes == kSlicedKernel) {
        if (primal) launch_resident(b, mult, k_e_half<true, false, VT>, 0, A, work, wc, i);
        else launch_resident(b, mult, k_e_half<false, false, VT>, 0, A, work, wc, i);
    } else if (lanes == kStagedKernel) {
        const int v = primal ? b->win_at : b->win_a;
        if (v > kStagedMaxVec) fail(PM_ERR_INVALID_ARG, "a column window is too wide for the staged iterate kernel");
        const size_t lds = sizeof(float) * static_cast<size_t>(std::max(v, 1));
        if (primal) launch_resident(b, mult, k_e_half<true, true, VT>, lds, A, work, wc, i);
        else launch_resident(b, mult, k_e_half<false, true, VT>, lds, A, work, wc, i);
    } else {
        with_lanes(lanes, [&](auto LC) {
            if (primal) launch_resident(b, mult, k_s_primal<decltype(LC)::value, VT>, 0, A, work, wc, i);
            else launch_resident(b, mult, k_s_dual<decltype(LC)::value, VT>, 0, A, work, wc, i);
        });
    }
    PM_LAUNCHED();
}

// Half `which` of iteration i with the kernels of the launches being issued:
// one launch over every LP, or one per size class.
template <class VT>
void launch_half(pm_batch* b, const BatchArrays<VT>& A, int which, int i) {
    const bool dual = which == PM_KERNEL_SPMV_A;
    if (!b->mode_bucketed) {
        launch_iterate(b, A, which, i, dual ? b->mode_a : b->mode_at, dual ? b->mode_mult_a : b->mode_mult_at);
        return;
    }
    for (int c = 0; c < pm_batch::kBuckets; ++c)
        if (b->bucket_lps[c] > 0)
            launch_iterate(b, A, which, i, dual ? b->blanes_a[c] : b->blanes_at[c],
                           dual ? b->bmult_a[c] : b->bmult_at[c], c);
}

// Work-list codes of a kernel: rows per item of a lanes kernel, or the code.
inline int items_code(int lanes) { return lanes > 0 ? kBlock / lanes : lanes; }

// Work list of size class `bucket` (-1: every LP, with the check blocks;
// -2: the check blocks only) for the given kernels.
template <class VT>
void launch_worklist(pm_batch* b, const BatchArrays<VT>& A, int bucket, int lanes_a, int lanes_at) {
    WorkItem* wr = b->work_r.get();
    WorkItem* wc = b->work_c.get();
    int* cnt = b->wcnt.get();
    if (bucket >= 0) {
        wr += b->off_r[bucket];
        wc += b->off_c[bucket];
        cnt += 4 + 2 * bucket;
    }
    hipLaunchKernelGGL(k_b_worklist<VT>, dim3(1), dim3(1024), 0, b->stream, A, b->count, bucket, items_code(lanes_a),
                       items_code(lanes_at), b->seg_a.get(), b->seg_at.get(), wr, wc, cnt);
    PM_LAUNCHED();
}

template <class VT>
void launch_phase(pm_batch* b, int phase, const VT* oa, const VT* oat) {
    hipStream_t s = b->stream;
    auto A = batch_arrays<VT>(b, oa, oat);
    BatchState* S = reinterpret_cast<BatchState*>(b->state.get());
    const dim3 g(b->count), blk(kBB);
    switch (phase) {
        case 0: hipLaunchKernelGGL(k_b_init<VT>, g, blk, 0, s, A, S); break;
        case 1:
            launch_resident(b, 1, k_b_items<kPassRound, VT>, 0, A, S);
            break;
        case 2:
            for (int i = 0; i < b->check_frequency; ++i) {
                launch_half(b, A, PM_KERNEL_SPMV_AT, i);
                launch_half(b, A, PM_KERNEL_SPMV_A, i);
            }
            return;
        case 4:
            if (!b->mode_bucketed) {
                launch_worklist(b, A, -1, b->mode_a, b->mode_at);
            } else {
                launch_worklist(b, A, -2, b->mode_a, b->mode_at);
                for (int c = 0; c < pm_batch::kBuckets; ++c)
                    if (b->bucket_lps[c] > 0) launch_worklist(b, A, c, b->blanes_a[c], b->blanes_at[c]);
            }
            return;
        case 3:
            launch_resident(b, 1, k_b_items<kPassTotals, VT>, 0, A, S);
            PM_LAUNCHED();
            launch_resident(b, 1, k_b_items<kPassKkt, VT>, 0, A, S);
            PM_LAUNCHED();
            hipLaunchKernelGGL(k_b_check<VT>, g, blk, 0, s, A, S, b->counter.get());
            break;
        default: fail(PM_ERR_INTERNAL, "unknown batch phase");
    }
    PM_LAUNCHED();
}

void phase(pm_batch* b, int p) {
    if (b->exact) launch_phase<float>(b, p, b->oa32.get(), b->oat32.get());
    else launch_phase<double>(b, p, b->oa64.get(), b->oat64.get());
}

// Kernels of the launches issued next: the tuned ones, or the tail ones.
void set_mode(pm_batch* b, bool tail) {
    b->mode_bucketed = !tail && b->nbuckets > 1;
    b->mode_a = tail ? b->tail_a : b->lanes_a;
    b->mode_at = tail ? b->tail_at : b->lanes_at;
    b->mode_mult_a = tail ? 1 : b->mult_a;
    b->mode_mult_at = tail ? 1 : b->mult_at;
}

// One chunk: clear the active counter, round data, K iterations, check.
void chunk_launches(pm_batch* b) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, b->stream, b->counter.get());
    PM_LAUNCHED();
    phase(b, 1);
    phase(b, 2);
    phase(b, 3);
    phase(b, 4);
}

void batch_run(pm_batch* b, pm_pdhg_result* results, pm_batch_stats* stats) {
    hipStream_t s = b->stream;
    // Tail kernels: taken once the running LPs hold at most tail_nnz
    // nonzeros (estimated from the running count and the mean LP), by a
    // second graph that first rebuilds the work list for them.
    const bool can_tail = b->tail_nnz > 0 && (b->tail_a != b->lanes_a || b->tail_at != b->lanes_at);
    const double mean_nnz = static_cast<double>(b->nnz_total) / b->count;
    if (b->use_graph && !b->exec) {
        set_mode(b, false);
        PM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < b->chunks_per_graph; ++q) chunk_launches(b);
        PM_HIP(hipStreamEndCapture(s, &b->graph));
        PM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        if (can_tail) {
            set_mode(b, true);
            PM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            phase(b, 4);
            for (int q = 0; q < b->chunks_per_graph; ++q) chunk_launches(b);
            PM_HIP(hipStreamEndCapture(s, &b->graph_tail));
            PM_HIP(hipGraphInstantiate(&b->exec_tail, b->graph_tail, nullptr, nullptr, 0));
        }
    }
    // Warm-start flags may have changed since the last run.
    b->d_desc.upload(b->desc, s);
    PM_HIP(hipEventRecord(b->ev0, s));
    set_mode(b, false);
    phase(b, 0);
    phase(b, 4);
    bool tail = false;
    const int max_launches = b->max_iterations / std::max(1, b->check_frequency) + 4 * b->max_refinements + 8;
    for (int q = 0; q < max_launches; q += b->chunks_per_graph) {
        if (b->use_graph) {
            PM_HIP(hipGraphLaunch(tail ? b->exec_tail : b->exec, s));
        } else {
            if (tail) phase(b, 4);
            for (int c = 0; c < b->chunks_per_graph; ++c) chunk_launches(b);
        }
        PM_HIP(hipMemcpyAsync(b->h_counter.get(), b->counter.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        PM_HIP(hipStreamSynchronize(s));
        if (b->h_counter[0] == 0) break;
        if (can_tail && !tail && b->h_counter[0] * mean_nnz <= static_cast<double>(b->tail_nnz)) {
            tail = true;
            set_mode(b, true);
        }
    }
    set_mode(b, false);
    PM_HIP(hipEventRecord(b->ev1, s));
    PM_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    std::vector<BatchDesc> out(b->count);
    b->d_desc.download(out.data(), b->count, s);
    PM_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < b->count; ++k) {
        b->desc[k].eta = out[k].eta;  // measured once, reused by later runs
        b->desc[k].warm_accepted = out[k].warm_accepted;
        b->desc[k].start_kkt = out[k].start_kkt;
    }
    if (stats) {
        stats->device_seconds = ms * 1e-3;
        if (stats->instance_seconds)
            for (int k = 0; k < b->count; ++k)
                stats->instance_seconds[k] =
                    static_cast<double>(out[k].t_end - out[k].t_start) / (b->clock_khz * 1e3);
    }
    if (!results) return;
    for (int k = 0; k < b->count; ++k) {
        const BatchDesc& d = out[k];
        pm_pdhg_result& r = results[k];
        r.status = d.status;
        r.iterations = d.iterations;
        r.restarts = d.restarts;
        r.refinements = d.refinements;
        r.polish_attempts = d.polishes;
        r.rel_kkt = d.rel_kkt;
        r.rel_primal = d.rel_p;
        r.rel_dual = d.rel_d;
        r.rel_gap = d.rel_g;
        r.primal_objective = d.pobj;
        r.dual_objective = d.dobj;
        r.primal_weight = d.omega;
        r.device_seconds = static_cast<double>(d.t_end - d.t_start) / (b->clock_khz * 1e3);
        r.primal_residual = NAN;
        r.dual_residual = NAN;
        const double sg = b->sense[k] == 1 ? -1.0 : 1.0;
        r.objective = sg * d.pobj + b->offset[k];
        if (r.x) PM_HIP(hipMemcpyAsync(r.x, b->xtot.get() + d.col_off, sizeof(double) * d.n, hipMemcpyDeviceToHost, s));
        if (r.y && d.m > 0)
            PM_HIP(hipMemcpyAsync(r.y, b->ytot.get() + d.row_off, sizeof(double) * d.m, hipMemcpyDeviceToHost, s));
        if (r.row_activity) std::fill(r.row_activity, r.row_activity + d.m, NAN);
    }
    PM_HIP(hipStreamSynchronize(s));
}

// Mean duration of one grid-wide iterate launch with every LP running.
void batch_time_kernel(pm_batch* b, int which, int reps, double* avg_ms, double* alg_bytes) {
    if (which != PM_KERNEL_SPMV_A && which != PM_KERNEL_SPMV_AT) fail(PM_ERR_INVALID_ARG, "unknown kernel id");
    if (reps <= 0) fail(PM_ERR_INVALID_ARG, "reps must be positive");
    hipStream_t s = b->stream;
    b->d_desc.upload(b->desc, s);
    set_mode(b, false);
    phase(b, 0);  // every LP active, chunk = check_frequency
    phase(b, 4);
    auto launch = [&](const auto& A) { launch_half(b, A, which, 0); };
    auto one = [&] {
        if (b->exact) launch(batch_arrays<float>(b, b->oa32.get(), b->oat32.get()));
        else launch(batch_arrays<double>(b, b->oa64.get(), b->oat64.get()));
    };
    one();  // warm caches and code
    PM_HIP(hipEventRecord(b->ev0, s));
    for (int r = 0; r < reps; ++r) one();
    PM_HIP(hipEventRecord(b->ev1, s));
    PM_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    double bytes = 0.0;
    for (const BatchDesc& d : b->desc) {
        const double nnz = d.nnz, m = d.m, n = d.n;
        bytes += which == PM_KERNEL_SPMV_A ? 8.0 * nnz + 4.0 * (m + 1) + 4.0 * n + 20.0 * m
                                           : 8.0 * nnz + 4.0 * (n + 1) + 4.0 * m + 28.0 * n;
    }
    if (avg_ms) *avg_ms = ms / reps;
    if (alg_bytes) *alg_bytes = bytes;
}

// Kernel of each half of the iteration (staged, sliced, stream, or 1..32
// lanes per row; the sliced ones on one or four resident grids): each
// candidate timed over the whole batch with every LP running (iterates
// zeroed, small steps), fastest kept.  Then, when the batch holds LPs of
// more than one size class, each class's fastest kernel is found on the
// class alone, and one launch per class is kept when the whole sequence of
// class launches (both halves, every class in turn, as a chunk issues
// them) takes under 97 % of the single launches' time.  A class timed
// alone can sit in the last-level cache across repeats, so only the whole
// sequence decides.  Runs once at creation.
void batch_autotune(pm_batch* b) {
    hipStream_t s = b->stream;
    for (auto* p : {&b->x, &b->xe, &b->x0, &b->xt, &b->xl, &b->c32, &b->lx32, &b->ux32})
        PM_HIP(hipMemsetAsync(p->get(), 0, sizeof(float) * p->size(), s));
    for (auto* p : {&b->y, &b->y0, &b->yt, &b->yl, &b->lr32, &b->ur32, &b->g32})
        PM_HIP(hipMemsetAsync(p->get(), 0, sizeof(float) * p->size(), s));
    std::vector<LpIter> it(b->count);
    for (int k = 0; k < b->count; ++k) {
        it[k] = LpIter{1e-3f, 1e-3f, 0, 1, 0, static_cast<int32_t>(b->desc[k].row_off),
                       static_cast<int32_t>(b->desc[k].col_off), 0};
    }
    b->iter.upload(it, s);
    b->nbuckets = 1;
    const int cand[] = {kStagedKernel, kSlicedKernel, kStreamKernel, 1, 2, 4, 8, 16, 32};
    // Mean time of one launch of `which` with kernel l on g grids over the
    // LPs of size class `bucket` (-1: every LP).
    auto time_launch = [&](int which, int l, int g, int bucket) {
        auto run = [&](const auto& A) {
            launch_worklist(b, A, bucket, l, l);
            launch_iterate(b, A, which, 0, l, g, bucket);
            PM_HIP(hipEventRecord(b->ev0, s));
            for (int r = 0; r < 3; ++r) launch_iterate(b, A, which, 0, l, g, bucket);
            PM_HIP(hipEventRecord(b->ev1, s));
        };
        if (b->exact) run(batch_arrays<float>(b, b->oa32.get(), b->oat32.get()));
        else run(batch_arrays<double>(b, b->oa64.get(), b->oat64.get()));
        PM_HIP(hipEventSynchronize(b->ev1));
        float ms = 0.0f;
        PM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        return static_cast<double>(ms) / 3.0;
    };
    auto best_of = [&](int which, int bucket, int& lanes, int& mult) {
        const int widest = which == PM_KERNEL_SPMV_A ? b->win_a : b->win_at;
        double best = INFINITY;
        for (int l : cand) {
            if ((l == kStagedKernel && widest > kStagedMaxVec) || (l == kSlicedKernel && widest > kSlicedMaxWin))
                continue;
            for (int g : {1, 4}) {
                if (g > 1 && l != kStagedKernel && l != kSlicedKernel) continue;
                const double ms = time_launch(which, l, g, bucket);
                if (ms < best) {
                    best = ms;
                    lanes = l;
                    mult = g;
                }
            }
        }
        return best;
    };
    double whole = 0.0, split = 0.0;  // sums of the best single-launch times, for the record
    for (int which : {PM_KERNEL_SPMV_AT, PM_KERNEL_SPMV_A}) {
        if (which == PM_KERNEL_SPMV_A && b->rows_total == 0) continue;
        const bool dual = which == PM_KERNEL_SPMV_A;
        whole += best_of(which, -1, dual ? b->lanes_a : b->lanes_at, dual ? b->mult_a : b->mult_at);
    }
    int classes = 0;
    for (int c = 0; c < pm_batch::kBuckets; ++c) {
        classes += b->bucket_lps[c] > 0;
        b->blanes_a[c] = b->lanes_a;
        b->blanes_at[c] = b->lanes_at;
        b->bmult_a[c] = b->mult_a;
        b->bmult_at[c] = b->mult_at;
    }
    if (classes > 1) {
        for (int which : {PM_KERNEL_SPMV_AT, PM_KERNEL_SPMV_A}) {
            if (which == PM_KERNEL_SPMV_A && b->rows_total == 0) continue;
            const bool dual = which == PM_KERNEL_SPMV_A;
            for (int c = 0; c < pm_batch::kBuckets; ++c)
                if (b->bucket_lps[c] > 0)
                    split += best_of(which, c, dual ? b->blanes_a[c] : b->blanes_at[c],
                                     dual ? b->bmult_a[c] : b->bmult_at[c]);
        }
        (void)split;
        // both halves as a chunk issues them, one launch over every LP
        // against one launch per class
        auto time_halves = [&](int nb) {
            b->nbuckets = nb;
            set_mode(b, false);
            auto run = [&](const auto& A) {
                if (nb > 1) {
                    launch_worklist(b, A, -2, b->lanes_a, b->lanes_at);
                    for (int c = 0; c < pm_batch::kBuckets; ++c)
                        if (b->bucket_lps[c] > 0) launch_worklist(b, A, c, b->blanes_a[c], b->blanes_at[c]);
                } else {
                    launch_worklist(b, A, -1, b->lanes_a, b->lanes_at);
                }
                launch_half(b, A, PM_KERNEL_SPMV_AT, 0);
                launch_half(b, A, PM_KERNEL_SPMV_A, 0);
                PM_HIP(hipEventRecord(b->ev0, s));
                for (int r = 0; r < 3; ++r) {
                    launch_half(b, A, PM_KERNEL_SPMV_AT, 0);
                    launch_half(b, A, PM_KERNEL_SPMV_A, 0);
                }
                PM_HIP(hipEventRecord(b->ev1, s));
            };
            if (b->exact) run(batch_arrays<float>(b, b->oa32.get(), b->oat32.get()));
            else run(batch_arrays<double>(b, b->oa64.get(), b->oat64.get()));
            PM_HIP(hipEventSynchronize(b->ev1));
            float ms = 0.0f;
            PM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            return static_cast<double>(ms) / 3.0;
        };
        const double t_one = time_halves(1), t_classes = time_halves(pm_batch::kBuckets);
        b->nbuckets = t_classes < 0.97 * t_one ? pm_batch::kBuckets : 1;
    }
    set_mode(b, false);
}

}  // namespace
}  // namespace pm

extern "C" {

int pm_batch_spmv_variants(pm_batch* b, int32_t* variant_a, int32_t* variant_at) {
    return pm::guarded([&] {
        if (!b || !variant_a || !variant_at) pm::fail(PM_ERR_INVALID_ARG, "NULL argument");
        *variant_a = pm::lanes_variant(b->lanes_a);
        *variant_at = pm::lanes_variant(b->lanes_at);
    });
}

int pm_batch_size_classes(pm_batch* b, int32_t* launches, int32_t* lps, int32_t* variant_a, int32_t* variant_at) {
    static_assert(pm_batch::kBuckets == PM_BATCH_SIZE_CLASSES, "size classes of the C-ABI");
    return pm::guarded([&] {
        if (!b) pm::fail(PM_ERR_INVALID_ARG, "batch is NULL");
        if (launches) *launches = b->nbuckets;
        for (int c = 0; c < pm_batch::kBuckets; ++c) {
            if (lps) lps[c] = b->bucket_lps[c];
            if (variant_a) variant_a[c] = pm::lanes_variant(b->nbuckets > 1 ? b->blanes_a[c] : b->lanes_a);
            if (variant_at) variant_at[c] = pm::lanes_variant(b->nbuckets > 1 ? b->blanes_at[c] : b->lanes_at);
        }
    });
}

int pm_batch_warm_start_verdicts(pm_batch* b, int32_t* accepted, double* start_rel_kkt) {
    return pm::guarded([&] {
        if (!b) pm::fail(PM_ERR_INVALID_ARG, "batch is NULL");
        for (int k = 0; k < b->count; ++k) {
            if (accepted) accepted[k] = b->desc[k].warm_accepted;
            if (start_rel_kkt) start_rel_kkt[k] = b->desc[k].start_kkt;
        }
    });
}

int pm_batch_set_size_classes(pm_batch* b, int32_t launches) {
    return pm::guarded([&] {
        if (!b) pm::fail(PM_ERR_INVALID_ARG, "batch is NULL");
        if (launches != 1 && launches != PM_BATCH_SIZE_CLASSES)
            pm::fail(PM_ERR_INVALID_ARG, "launches must be 1 or PM_BATCH_SIZE_CLASSES");
        PM_HIP(hipSetDevice(b->device));
        PM_HIP(hipStreamSynchronize(b->stream));
        b->nbuckets = launches;
        pm::set_mode(b, false);
        // the captured graphs issue the previous launches: capture again
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->exec_tail) (void)hipGraphExecDestroy(b->exec_tail);
        if (b->graph_tail) (void)hipGraphDestroy(b->graph_tail);
        b->exec = b->exec_tail = nullptr;
        b->graph = b->graph_tail = nullptr;
    });
}

int pm_batch_set_graph(pm_batch* b, int32_t use_graph) {
    return pm::guarded([&] {
        if (!b) pm::fail(PM_ERR_INVALID_ARG, "batch is NULL");
        b->use_graph = use_graph != 0;
    });
}

int pm_batch_time_kernel(pm_batch* b, int32_t which, int32_t reps, double* avg_ms, double* alg_bytes) {
    return pm::guarded([&] {
        if (!b) pm::fail(PM_ERR_INVALID_ARG, "batch is NULL");
        PM_HIP(hipSetDevice(b->device));
        pm::batch_time_kernel(b, which, reps, avg_ms, alg_bytes);
    });
}

int pm_batch_create(int32_t device, int32_t count, const pm_lp* lps, const pm_pdhg_options* opts, pm_batch** out) {
    pm::Range range_("pm_batch_create");
    return pm::guarded([&] {
        if (!out) pm::fail(PM_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        int nd = 0;
        if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
            (void)hipGetLastError();
            pm::fail(PM_ERR_NO_DEVICE, "no HIP device visible");
        }
        if (device < 0 || device >= nd) pm::fail(PM_ERR_NO_DEVICE, "device ordinal out of range");
        PM_HIP(hipSetDevice(device));
        auto b = std::make_unique<pm_batch>();
        b->device = device;
        PM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        PM_HIP(hipEventCreate(&b->ev0));
        PM_HIP(hipEventCreate(&b->ev1));
        pm::batch_create(b.get(), count, lps, opts);
        *out = b.release();
    });
}

int pm_batch_set_warm_start(pm_batch* b, int32_t i, const double* x0, const double* y0) {
    return pm::guarded([&] {
        if (!b || i < 0 || i >= b->count) pm::fail(PM_ERR_INVALID_ARG, "bad batch or index");
        PM_HIP(hipSetDevice(b->device));
        pm::BatchDesc& d = b->desc[i];
        d.warm_x = x0 != nullptr;
        d.warm_y = y0 != nullptr && d.m > 0;
        d.omega_scale = 0.0;  // a learned primal weight belongs to the learned start
        if (x0) PM_HIP(hipMemcpyAsync(b->wx.get() + d.col_off, x0, sizeof(double) * d.n, hipMemcpyHostToDevice, b->stream));
        if (y0 && d.m > 0)
            PM_HIP(hipMemcpyAsync(b->wy.get() + d.row_off, y0, sizeof(double) * d.m, hipMemcpyHostToDevice, b->stream));
        PM_HIP(hipStreamSynchronize(b->stream));
    });
}

int pm_batch_set_primal_weight(pm_batch* b, int32_t i, double omega) {
    return pm::guarded([&] {
        if (!b || i < 0 || i >= b->count) pm::fail(PM_ERR_INVALID_ARG, "bad batch or index");
        if (!std::isfinite(omega)) pm::fail(PM_ERR_INVALID_ARG, "primal weight must be finite");
        b->desc[i].omega0 = omega;
    });
}

int pm_batch_scaling(pm_batch* b, int32_t i, double* row_scale, double* col_scale) {
    return pm::guarded([&] {
        if (!b || i < 0 || i >= b->count) pm::fail(PM_ERR_INVALID_ARG, "bad batch or index");
        if (!col_scale || (!row_scale && b->desc[i].m > 0)) pm::fail(PM_ERR_INVALID_ARG, "NULL output");
        PM_HIP(hipSetDevice(b->device));
        const pm::BatchDesc& d = b->desc[i];
        if (d.m > 0)
            PM_HIP(hipMemcpyAsync(row_scale, b->dr.get() + d.row_off, sizeof(double) * d.m, hipMemcpyDeviceToHost,
                                  b->stream));
        PM_HIP(hipMemcpyAsync(col_scale, b->dc.get() + d.col_off, sizeof(double) * d.n, hipMemcpyDeviceToHost,
                              b->stream));
        PM_HIP(hipStreamSynchronize(b->stream));
    });
}

int pm_batch_run(pm_batch* b, pm_pdhg_result* results, pm_batch_stats* stats) {
    pm::Range range_("pm_batch_run");
    return pm::guarded([&] {
        if (!b) pm::fail(PM_ERR_INVALID_ARG, "batch is NULL");
        PM_HIP(hipSetDevice(b->device));
        pm::batch_run(b, results, stats);
    });
}

int pm_batch_destroy(pm_batch* b) {
    return pm::guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->exec_tail) (void)hipGraphExecDestroy(b->exec_tail);
        if (b->graph_tail) (void)hipGraphDestroy(b->graph_tail);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        hipStream_t s = b->stream;
        delete b;
        if (s) (void)hipStreamDestroy(s);
    });
}

int pm_batch_solve(int32_t count, const pm_lp* lps, const double* const* x0, const double* const* y0,
                   const pm_pdhg_options* opts, pm_pdhg_result* results, pm_batch_stats* stats) {
    pm_batch* b = nullptr;
    int rc = pm_batch_create(opts ? opts->device : 0, count, lps, opts, &b);
    if (rc != PM_OK) return rc;
    for (int k = 0; k < count && rc == PM_OK; ++k)
        if ((x0 && x0[k]) || (y0 && y0[k])) rc = pm_batch_set_warm_start(b, k, x0 ? x0[k] : nullptr, y0 ? y0[k] : nullptr);
    if (rc == PM_OK) rc = pm_batch_run(b, results, stats);
    const std::string err = rc != PM_OK ? std::string(pm_last_error()) : std::string();
    pm_batch_destroy(b);
    if (rc != PM_OK) pm::set_last_error(err);
    return rc;
}

}  // extern "C"
