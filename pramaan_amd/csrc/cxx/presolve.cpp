// presolve.cpp -- pramaan::ruiz_scale and unscale_solution over the C-ABI
// (reference src/presolve/scaling.cpp:29-187).  The passes run on the
// device (pm_ruiz_scale); this file moves the model's arrays in and out and
// keeps the reference's contract: the model is scaled in place, its CSR
// rebuilt with the scaled values, one CumulativeScaling appended to the
// ledger, the factors returned.
#include <stdexcept>
#include <string>
#include <utility>

#include "pramaan/presolve.hpp"
#include "pramaan_hip.h"

namespace pramaan {

ScalingFactors ruiz_scale(ModelIR& model, TransformationLedger& ledger, int max_iterations) {
    const int m = model.numRows();
    const int n = model.numVars();
    ScalingFactors f;
    f.row_scale.assign(static_cast<std::size_t>(m), 1.0);
    f.col_scale.assign(static_cast<std::size_t>(n), 1.0);
    // scaling.cpp:37 returns before touching the model or the ledger
    if (m == 0 || n == 0) return f;
    std::vector<CSRMatrix::Index> row_ptr = model.A.rowPtr();
    std::vector<CSRMatrix::Index> col_idx = model.A.colIdx();
    std::vector<double> values = model.A.values();
    const int rc = pm_ruiz_scale(m, n, row_ptr.data(), col_idx.data(), values.data(), model.obj_coeffs.data(),
                                 model.row_lower.data(), model.row_upper.data(), model.var_lower.data(),
                                 model.var_upper.data(), max_iterations > 0 ? max_iterations : 0,
                                 f.row_scale.data(), f.col_scale.data());
    if (rc != PM_OK) throw std::runtime_error(std::string("HIP Ruiz scaling error: ") + pm_last_error());
    model.A = CSRMatrix(std::move(row_ptr), std::move(col_idx), std::move(values), n);
    CumulativeScaling rec;
    rec.row_scale = f.row_scale;
    rec.col_scale = f.col_scale;
    ledger.record(std::move(rec));
    return f;
}

void unscale_solution(std::vector<double>& x, double& objective_value, const ModelIR& pre_scaled_model,
                      const ScalingFactors& factors) {
    const std::size_t n = x.size();
    for (std::size_t j = 0; j < n; ++j) x[j] *= factors.col_scale[j];
    double obj = pre_scaled_model.obj_offset;
    for (std::size_t j = 0; j < n; ++j) obj += pre_scaled_model.obj_coeffs[j] * x[j];
    objective_value = obj;
}

}  // namespace pramaan
