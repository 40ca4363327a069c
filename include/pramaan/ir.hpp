// ir.hpp -- the LP/MILP model of the PRAMAAN interface (ModelIR).
//
// Same fields, constructor and helpers as the reference's ModelIR
// (include/pramaan/ir.hpp:28-123): ranged rows row_lower <= A x <= row_upper,
// boxed variables, objective sense and offset, names and variable types;
// +-kInfinity = 1e30 marks an absent bound.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pramaan/sparse_matrix.hpp"

namespace pramaan {

enum class ObjSense { kMinimize, kMaximize };
enum class VarType { kContinuous, kInteger };

inline constexpr double kInfinity = 1.0e30;

struct ModelIR {
    ModelIR() = default;
    ModelIR(ObjSense obj_sense_in, double obj_offset_in, std::vector<double> obj_coeffs_in, CSRMatrix A_in,
            std::vector<double> row_lower_in, std::vector<double> row_upper_in, std::vector<std::string> row_names_in,
            std::vector<double> var_lower_in, std::vector<double> var_upper_in, std::vector<VarType> var_types_in,
            std::vector<std::string> var_names_in);

    ObjSense obj_sense = ObjSense::kMinimize;
    double obj_offset = 0.0;
    std::vector<double> obj_coeffs;
    CSRMatrix A;
    std::vector<double> row_lower;
    std::vector<double> row_upper;
    std::vector<std::string> row_names;
    std::vector<double> var_lower;
    std::vector<double> var_upper;
    std::vector<VarType> var_types;
    std::vector<std::string> var_names;

    CSRMatrix::Index numVars() const noexcept { return A.numCols(); }
    CSRMatrix::Index numRows() const noexcept { return A.numRows(); }
    bool isInteger(CSRMatrix::Index j) const { return var_types[j] == VarType::kInteger; }
    bool isPureLP() const {
        for (VarType t : var_types)
            if (t == VarType::kInteger) return false;
        return true;
    }
    bool isEqualityRow(CSRMatrix::Index r) const { return row_lower[r] == row_upper[r]; }
    bool isFreeRow(CSRMatrix::Index r) const { return row_lower[r] <= -kInfinity && row_upper[r] >= kInfinity; }

    // Throws std::invalid_argument naming the first broken invariant.
    void validate() const;
};

std::uint64_t computeModelFingerprint(const ModelIR& model);
std::uint64_t computeStructuralFingerprint(const ModelIR& model);

}  // namespace pramaan
