// dispatch.hpp -- host-side launch of the templated row kernels by SpMV
// variant (PM_SPMV_*), and the small elementwise/reduction launchers.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace pm {

template <int V>
using IntC = std::integral_constant<int, V>;
template <bool V>
using BoolC = std::integral_constant<bool, V>;

// Lanes per row of a variant (the LDS-stream id falls back to 8 lanes).
inline int variant_lanes(int variant) {
    switch (variant) {
        case PM_SPMV_SEQUENTIAL:
        case PM_SPMV_SCALAR: return 1;
        case PM_SPMV_VECTOR2: return 2;
        case PM_SPMV_VECTOR4: return 4;
        case PM_SPMV_VECTOR8: return 8;
        case PM_SPMV_VECTOR16: return 16;
        case PM_SPMV_VECTOR32: return 32;
        case PM_SPMV_VECTOR64: return 64;
        default: return 8;
    }
}

inline int grid_for(int nrows, int lanes) {
    const long long threads = static_cast<long long>(nrows) * lanes;
    return static_cast<int>((threads + kBlock - 1) / kBlock);
}

template <class F>
void with_variant(int variant, F&& f) {
    switch (variant) {
        case PM_SPMV_SEQUENTIAL: f(IntC<1>{}, BoolC<true>{}); break;
        case PM_SPMV_SCALAR: f(IntC<1>{}, BoolC<false>{}); break;
        case PM_SPMV_VECTOR2: f(IntC<2>{}, BoolC<false>{}); break;
        case PM_SPMV_VECTOR4: f(IntC<4>{}, BoolC<false>{}); break;
        case PM_SPMV_VECTOR16: f(IntC<16>{}, BoolC<false>{}); break;
        case PM_SPMV_VECTOR32: f(IntC<32>{}, BoolC<false>{}); break;
        case PM_SPMV_VECTOR64: f(IntC<64>{}, BoolC<false>{}); break;
        case PM_SPMV_VECTOR8:
        default: f(IntC<8>{}, BoolC<false>{}); break;
    }
}

// Lanes per row as a compile-time constant (1, 2, ..., 32).
template <class F>
void with_lanes(int lanes, F&& f) {
    switch (lanes) {
        case 1: f(IntC<1>{}); break;
        case 2: f(IntC<2>{}); break;
        case 4: f(IntC<4>{}); break;
        case 16: f(IntC<16>{}); break;
        case 32: f(IntC<32>{}); break;
        case 8:
        default: f(IntC<8>{}); break;
    }
}

// y-side row kernel over `nrows` rows.
template <class AT, class VT, class XT, class Epi>
void launch_rows(int variant, int nrows, const int* rp, const int* ci, const VT* v, const XT* x,
                 const Epi& epi, hipStream_t s) {
    if (nrows <= 0) return;
    with_variant(variant, [&](auto Lc, auto Ec) {
        constexpr int L = decltype(Lc)::value;
        constexpr bool E = decltype(Ec)::value;
        hipLaunchKernelGGL((k_rows<L, E, AT, VT, XT, Epi>), dim3(grid_for(nrows, L)), dim3(kBlock), 0, s,
                           nrows, rp, ci, v, x, epi);
    });
    PM_LAUNCHED();
}

// Sliced ELL product over nslices slices of 64 rows (k_sell).
template <class Epi>
void launch_sell(int nslices, const int* sptr, const int* slen, const int* srow, const int* col, const float* val,
                 const float* x, const Epi& epi, hipStream_t s) {
    if (nslices <= 0) return;
    const int per = kBlock / 64;
    hipLaunchKernelGGL((k_sell<Epi>), dim3((nslices + per - 1) / per), dim3(kBlock), 0, s, nslices, sptr, slen, srow,
                       col, val, x, epi);
    PM_LAUNCHED();
}

// Row kernel with block-partial reduction, followed by the final reduce of
// its partials into out[0 .. NS+NM).  `partials` must hold
// grid_for(nrows, lanes) * (NS+NM) doubles.
template <class AT, class VT, class XT, class Epi>
void launch_rows_reduce(int variant, int nrows, const int* rp, const int* ci, const VT* v,
                        const XT* x, const Epi& epi, double* partials, double* out, hipStream_t s) {
    constexpr int NA = Epi::NS + Epi::NM;
    int grid = 0;
    if (nrows > 0) {
        with_variant(variant, [&](auto Lc, auto Ec) {
            constexpr int L = decltype(Lc)::value;
            constexpr bool E = decltype(Ec)::value;
            grid = grid_for(nrows, L);
            hipLaunchKernelGGL((k_rows_reduce<L, E, AT, VT, XT, Epi>), dim3(grid), dim3(kBlock), 0, s,
                               nrows, rp, ci, v, x, epi, partials);
        });
        PM_LAUNCHED();
    }
    hipLaunchKernelGGL(k_final_reduce<>, dim3(1), dim3(kBlock), 0, s, partials, grid, Epi::NS, Epi::NM, out);
    PM_LAUNCHED();
    (void)NA;
}

// Tiles of a merge-path product over nrows rows and nnz nonzeros (the size
// of its carry and head buffers; its partition holds one more).
inline int merge_tiles(int nrows, int nnz) {
    return static_cast<int>((static_cast<int64_t>(nrows) + nnz + kMergeTile - 1) / kMergeTile);
}

// Rows of a merge-path product that k_merge_fix finishes, from the host row
// pointers: more than kMergeShort entries and a tile edge inside the row
// (between its first nonzero's step rp[r] + r and its end step rp[r+1] + r).
inline std::vector<int4> merge_long_rows(int nrows, const int* rp) {
    std::vector<int4> out;
    for (int r = 0; r < nrows; ++r) {
        if (rp[r + 1] - rp[r] <= kMergeShort) continue;
        const int b0 = static_cast<int>((static_cast<int64_t>(rp[r]) + r) / kMergeTile);
        const int b1 = static_cast<int>((static_cast<int64_t>(rp[r + 1]) + r) / kMergeTile);
        if (b1 > b0) out.push_back(make_int4(r, b0, b1, 0));
    }
    return out;
}

// Tile starts of a merge-path product (prow: merge_tiles + 1 ints), once
// per matrix.
inline void launch_merge_partition(int nrows, int nnz, const int* rp, int* prow, hipStream_t s) {
    const int nt = merge_tiles(nrows, nnz);
    hipLaunchKernelGGL(k_merge_partition<kMergeTile>, dim3((nt + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nrows,
                       nnz, nt, rp, prow);
    PM_LAUNCHED();
}

// Merge-path product with a fused epilogue (k_merge_tile, then k_merge_fix
// when the matrix has long rows crossing tile edges) over a partition from
// launch_merge_partition and the list from merge_long_rows (nlong entries
// on the device); carry and head hold merge_tiles(nrows, nnz) floats.
template <class Epi>
void launch_merge(int nrows, int nnz, const int* rp, const int* ci, const float* v, const float* x, const Epi& epi,
                  const int* prow, const int4* longs, int nlong, float* carry, float* head, hipStream_t s) {
    if (nrows <= 0) return;
    const int nt = merge_tiles(nrows, nnz);
    hipLaunchKernelGGL((k_merge_tile<Epi>), dim3(nt), dim3(kBlock), 0, s, nrows, nnz, rp, ci, v, x, prow, carry, head,
                       epi);
    if (nlong > 0)
        hipLaunchKernelGGL((k_merge_fix<Epi>), dim3((nlong + kBlock / kWave - 1) / (kBlock / kWave)), dim3(kBlock), 0,
                           s, nlong, longs, carry, head, epi);
    PM_LAUNCHED();
}

template <class F>
void launch_elem(int n, const F& f, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_elem<F>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, f);
    PM_LAUNCHED();
}

template <class F>
void launch_elem_reduce(int n, const F& f, double* partials, double* out, hipStream_t s) {
    const int grid = n > 0 ? (n + kBlock - 1) / kBlock : 0;
    if (grid > 0) {
        hipLaunchKernelGGL((k_elem_reduce<F>), dim3(grid), dim3(kBlock), 0, s, n, f, partials);
        PM_LAUNCHED();
    }
    hipLaunchKernelGGL(k_final_reduce<>, dim3(1), dim3(kBlock), 0, s, partials, grid, F::NS, F::NM, out);
    PM_LAUNCHED();
}

// Partial-buffer size (doubles) needed by a reduce over `nrows` rows.
inline size_t partials_needed(int nrows, int width) {
    return static_cast<size_t>(grid_for(nrows > 0 ? nrows : 1, 64)) * static_cast<size_t>(width) + 64;
}

}  // namespace pm
