// gnn_features.cpp -- graph and features of one LP for the warm-start
// network (pramaan_gnn.h): the row / column max-norm scaled matrix in CSR
// and CSC order and the 16 column and 16 row features.  Host code; a few
// tenths of a millisecond per rolling-horizon window, so training batches
// can be rebuilt from the LP every step instead of being stored.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pramaan_gnn.h"
#include "pramaan_hip.h"

namespace {
inline bool fin(double v) { return std::fabs(v) < 1e29; }
// Bits of a node's index within its period as -1 / +1; all 0 for role < 0.
inline void role_bits(float* f, int role) {
    for (int b = 0; b < PM_GNN_ROLE_BITS; ++b) f[b] = role < 0 ? 0.0f : (((role >> b) & 1) ? 1.0f : -1.0f);
}
}  // namespace

extern "C" int pm_gnn_lp_features(const pm_lp* lp, int32_t period_cols, int32_t period_rows, int32_t* cp,
                                  int32_t* ri, float* ar, float* ac, float* xr, float* xc, double* scales) {
    if (!lp || !cp || !ri || !ar || !ac || !xr || !xc || !scales) return PM_ERR_INVALID_ARG;
    const int m = lp->m, n = lp->n, nnz = lp->nnz;
    if (m < 0 || n < 0 || nnz < 0 || !lp->row_ptr || (nnz > 0 && (!lp->col_idx || !lp->values))) return PM_ERR_INVALID_ARG;
    const int32_t* rp = lp->row_ptr;
    const int32_t* ci = lp->col_idx;
    const double* v = lp->values;
    for (int q = 0; q < nnz; ++q)
        if (ci[q] < 0 || ci[q] >= n) return PM_ERR_INVALID_ARG;
    // scales: s = max(1, largest finite |bound|), cs = max(1e-12, max |c|)
    double s = 0.0;
    auto upd = [&](const double* a, int k) {
        for (int i = 0; i < k; ++i)
            if (fin(a[i])) s = std::max(s, std::fabs(a[i]));
    };
    upd(lp->row_lower, m);
    upd(lp->row_upper, m);
    upd(lp->var_lower, n);
    upd(lp->var_upper, n);
    s = std::max(1.0, s);
    double cs = 0.0;
    for (int j = 0; j < n; ++j) cs = std::max(cs, std::fabs(lp->c[j]));
    cs = n > 0 ? std::max(1e-12, cs) : 1.0;
    scales[0] = s;
    scales[1] = cs;
    // row / column max-norms
    std::vector<double> rmax(m, 0.0), cmax(n, 0.0);
    std::vector<int32_t> collen(n, 0);
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const double a = std::fabs(v[k]);
            rmax[i] = std::max(rmax[i], a);
            cmax[ci[k]] = std::max(cmax[ci[k]], a);
            ++collen[ci[k]];
        }
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) ar[k] = static_cast<float>(v[k] / (rmax[i] > 0.0 ? rmax[i] : 1.0));
    // CSC by a stable counting sort: rows ascending inside each column
    cp[0] = 0;
    for (int j = 0; j < n; ++j) cp[j + 1] = cp[j] + collen[j];
    std::vector<int32_t> next(cp, cp + n);
    for (int i = 0; i < m; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k], q = next[j]++;
            ri[q] = i;
            ac[q] = static_cast<float>(v[k] / (cmax[j] > 0.0 ? cmax[j] : 1.0));
        }
    int maxc = 1, maxr = 1;
    for (int j = 0; j < n; ++j) maxc = std::max(maxc, collen[j]);
    for (int i = 0; i < m; ++i) maxr = std::max(maxr, rp[i + 1] - rp[i]);
    const int pc = period_cols > 0 ? std::max(1, n / period_cols - 1) : 1;
    const int pr = period_rows > 0 ? std::max(1, m / period_rows - 1) : 1;
    for (int j = 0; j < n; ++j) {
        const double lo = lp->var_lower[j], hi = lp->var_upper[j];
        float* f = xc + static_cast<size_t>(j) * PM_GNN_FEATURES;
        f[0] = static_cast<float>(lp->c[j] / cs);
        f[1] = static_cast<float>(fin(lo) ? lo / s : 0.0);
        f[2] = static_cast<float>(fin(hi) ? hi / s : 0.0);
        f[3] = lo > -1e29 ? 1.0f : 0.0f;
        f[4] = hi < 1e29 ? 1.0f : 0.0f;
        f[5] = static_cast<float>(static_cast<double>(collen[j]) / maxc);
        f[6] = period_cols > 0 ? static_cast<float>(static_cast<double>(j / period_cols) / pc) : 0.0f;
        role_bits(f + 7, period_cols > 0 ? j % period_cols : -1);
    }
    for (int i = 0; i < m; ++i) {
        const double lo = lp->row_lower[i], hi = lp->row_upper[i];
        float* f = xr + static_cast<size_t>(i) * PM_GNN_FEATURES;
        f[0] = static_cast<float>(fin(lo) ? lo / s : 0.0);
        f[1] = static_cast<float>(fin(hi) ? hi / s : 0.0);
        f[2] = lo > -1e29 ? 1.0f : 0.0f;
        f[3] = hi < 1e29 ? 1.0f : 0.0f;
        f[4] = lo == hi ? 1.0f : 0.0f;
        f[5] = static_cast<float>(static_cast<double>(rp[i + 1] - rp[i]) / maxr);
        f[6] = period_rows > 0 ? static_cast<float>(static_cast<double>(i / period_rows) / pr) : 0.0f;
        role_bits(f + 7, period_rows > 0 ? i % period_rows : -1);
    }
    return PM_OK;
}
