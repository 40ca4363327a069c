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

// Lanes per row for an average row length (about three entries per lane).
inline int lanes_for(double avg) {
    const int want = static_cast<int>(std::ceil(avg / 3.0));
    int l = 1;
    while (l < want && l < 32) l <<= 1;
    return l;
}

}  // namespace pm
