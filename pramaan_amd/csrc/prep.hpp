// prep.hpp -- host-side preparation of one LP for the device engines:
// counting-sort transpose and Ruiz + Pock-Chambolle diagonal scaling.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace pm {

// Transpose of an m x n CSR by counting sort over rows in increasing order,
// so each transposed row keeps its entries sorted (the reference's
// CSRMatrix::transpose, src/linalg/sparse_matrix.cpp:13-57).  perm[d] is the
// CSR position of transposed entry d.
inline void csr_transpose(int m, int n, const int* rp, const int* ci, std::vector<int>& cp, std::vector<int>& ri,
                          std::vector<int64_t>& perm) {
    const int nnz = rp[m];
    cp.assign(static_cast<size_t>(n) + 1, 0);
    for (int k = 0; k < nnz; ++k) ++cp[static_cast<size_t>(ci[k]) + 1];
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    std::vector<int> cur(cp.begin(), cp.end() - 1);
    ri.resize(nnz);
    perm.resize(nnz);
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int d = cur[ci[k]]++;
            ri[d] = i;
            perm[d] = k;
        }
}

// Ruiz equilibration (`passes` inf-norm passes over A) followed by one
// Pock-Chambolle alpha = 1 pass (1-norms): A_hat = D_r A D_c.  Returns the
// scaled values in CSR order.
inline std::vector<double> ruiz_pc_scale(int m, int n, const int* rp, const int* ci, const double* v, int passes,
                                         std::vector<double>& dr, std::vector<double>& dc) {
    const int nnz = rp[m];
    std::vector<double> vs(v, v + nnz), rn(m), cn(n);
    dr.assign(m, 1.0);
    dc.assign(n, 1.0);
    auto apply = [&]() {
        for (int i = 0; i < m; ++i) rn[i] = rn[i] > 0.0 ? 1.0 / std::sqrt(rn[i]) : 1.0;
        for (int j = 0; j < n; ++j) cn[j] = cn[j] > 0.0 ? 1.0 / std::sqrt(cn[j]) : 1.0;
        for (int i = 0; i < m; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) vs[k] *= rn[i] * cn[ci[k]];
        for (int i = 0; i < m; ++i) dr[i] *= rn[i];
        for (int j = 0; j < n; ++j) dc[j] *= cn[j];
    };
    for (int p = 0; p < passes; ++p) {
        std::fill(rn.begin(), rn.end(), 0.0);
        std::fill(cn.begin(), cn.end(), 0.0);
        for (int i = 0; i < m; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const double a = std::abs(vs[k]);
                rn[i] = std::max(rn[i], a);
                cn[ci[k]] = std::max(cn[ci[k]], a);
            }
        apply();
    }
    std::fill(rn.begin(), rn.end(), 0.0);
    std::fill(cn.begin(), cn.end(), 0.0);
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const double a = std::abs(vs[k]);
            rn[i] += a;
            cn[ci[k]] += a;
        }
    apply();
    return vs;
}

// Lanes per row for an average row length (about three entries per lane).
inline int lanes_for(double avg) {
    const int want = static_cast<int>(std::ceil(avg / 3.0));
    int l = 1;
    while (l < want && l < 32) l <<= 1;
    return l;
}

}  // namespace pm
