// transformation_ledger.hpp -- the append-only record of the reversible
// transformations applied to a model before it is solved (reference
// include/pramaan/transformation_ledger.hpp:17-86).  The record types and
// the ledger's interface are the reference's, so code that builds or reads
// a ledger compiles against either; ruiz_scale (presolve.hpp) appends one
// CumulativeScaling.  A certificate reports the number of entries.
#pragma once

#include <string>
#include <utility>
#include <variant>
#include <vector>

namespace pramaan {

// Variable `original_index` (lb == ub) was removed with value fixed_value;
// row_coeffs are its nonzeros, kept to check the right-hand-side shift.
struct FixedVariableRemoved {
    int original_index;
    double fixed_value;
    std::string var_name;
    double obj_coeff;
    struct RowCoeff {
        int row;
        double value;
    };
    std::vector<RowCoeff> row_coeffs;
};

// Row `row` was multiplied by `factor`.
struct RowScaled {
    int row;
    double factor;
};

// Column `col` was multiplied by `factor` (x_orig = x_scaled * factor).
struct ColumnScaled {
    int col;
    double factor;
};

// Every row r multiplied by row_scale[r] and column j by col_scale[j]
// (x_orig[j] = x_scaled[j] * col_scale[j]).
struct CumulativeScaling {
    std::vector<double> row_scale;
    std::vector<double> col_scale;
};

using Reduction = std::variant<FixedVariableRemoved, RowScaled, ColumnScaled, CumulativeScaling>;

class TransformationLedger {
public:
    void record(Reduction r) { reductions_.push_back(std::move(r)); }
    std::size_t size() const noexcept { return reductions_.size(); }
    bool empty() const noexcept { return reductions_.empty(); }
    const std::vector<Reduction>& reductions() const noexcept { return reductions_; }
    void clear() { reductions_.clear(); }

private:
    std::vector<Reduction> reductions_;
};

}  // namespace pramaan
