// presolve.hpp -- Ruiz equilibration of a model and its reversal (reference
// include/pramaan/presolve.hpp:33-54), the scaling on the device.
//
// ruiz_scale runs pm_ruiz_scale (pramaan_hip.h): max_iterations passes of
// row and column infinity norms, the column norms including |c_j|, rows
// and columns scaled by 1/sqrt(norm); A, c, the finite row bounds (times
// the row factor) and the finite variable bounds (divided by the column
// factor) are scaled in place and the cumulative factors are returned and
// recorded in the ledger as one CumulativeScaling.  The results equal the
// reference's host loops (src/presolve/scaling.cpp:29-150) bit for bit.
// It throws std::runtime_error with the library's message when the device
// call fails; there is no host fallback.
//
// unscale_solution maps x back (x_orig[j] = x_scaled[j] * col_scale[j]) and
// recomputes the objective from the unscaled model's offset and costs in
// index order (scaling.cpp:171-187).
//
// The reference's fixed-variable presolve (presolve_fixed_variables,
// postsolve) is a host-side model reduction outside this library.
#pragma once

#include <vector>

#include "pramaan/ir.hpp"
#include "pramaan/transformation_ledger.hpp"

namespace pramaan {

struct ScalingFactors {
    std::vector<double> row_scale;  // row r of A was multiplied by row_scale[r]
    std::vector<double> col_scale;  // column j of A was multiplied by col_scale[j]
};

ScalingFactors ruiz_scale(ModelIR& model, TransformationLedger& ledger, int max_iterations = 10);

void unscale_solution(std::vector<double>& x, double& objective_value, const ModelIR& pre_scaled_model,
                      const ScalingFactors& factors);

}  // namespace pramaan
