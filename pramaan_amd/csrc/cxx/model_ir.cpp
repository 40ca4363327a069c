// model_ir.cpp -- out-of-line parts of the model interface: ModelIR
// construction/validation, CSRMatrix::transpose and the two 64-bit FNV-1a
// model fingerprints.  The fingerprints hash the same fields in the same
// order as the reference (src/ir/model_ir.cpp:96-172), so a certificate's
// fingerprint written by either side verifies on the other.
#include <cstring>
#include <utility>

#include "pramaan/ir.hpp"

namespace pramaan {

ModelIR::ModelIR(ObjSense s, double off, std::vector<double> c, CSRMatrix A_in, std::vector<double> rl,
                 std::vector<double> ru, std::vector<std::string> rn, std::vector<double> vl,
                 std::vector<double> vu, std::vector<VarType> vt, std::vector<std::string> vn)
    : obj_sense(s), obj_offset(off), obj_coeffs(std::move(c)), A(std::move(A_in)), row_lower(std::move(rl)),
      row_upper(std::move(ru)), row_names(std::move(rn)), var_lower(std::move(vl)), var_upper(std::move(vu)),
      var_types(std::move(vt)), var_names(std::move(vn)) {
    validate();
}

void ModelIR::validate() const {
    // Every per-variable and per-row array must match the matrix, and no
    // lower bound may exceed its upper bound.
    const std::size_t n = static_cast<std::size_t>(numVars()), m = static_cast<std::size_t>(numRows());
    const auto length = [](const char* what, std::size_t got, std::size_t want, const char* per) {
        if (got != want)
            throw std::invalid_argument(std::string("ModelIR: ") + what + " holds " + std::to_string(got) +
                                        " entries for " + std::to_string(want) + " " + per);
    };
    length("the objective", obj_coeffs.size(), n, "columns");
    length("the variable lower bounds", var_lower.size(), n, "columns");
    length("the variable upper bounds", var_upper.size(), n, "columns");
    length("the variable types", var_types.size(), n, "columns");
    length("the variable names", var_names.size(), n, "columns");
    length("the row lower bounds", row_lower.size(), m, "rows");
    length("the row upper bounds", row_upper.size(), m, "rows");
    length("the row names", row_names.size(), m, "rows");
    const auto crossed = [](const std::vector<double>& lo, const std::vector<double>& hi, const char* what) {
        for (std::size_t k = 0; k < lo.size(); ++k)
            if (lo[k] > hi[k])
                throw std::invalid_argument(std::string("ModelIR: ") + what + " " + std::to_string(k) +
                                            " has its lower bound above its upper bound");
    };
    crossed(var_lower, var_upper, "variable");
    crossed(row_lower, row_upper, "row");
}

CSRMatrix CSRMatrix::transpose() const {
    std::vector<Index> tp(static_cast<std::size_t>(num_cols_) + 1, 0);
    for (Index c : col_idx_) ++tp[static_cast<std::size_t>(c) + 1];
    for (Index c = 0; c < num_cols_; ++c) tp[c + 1] += tp[c];
    std::vector<Index> ti(col_idx_.size()), cur(tp.begin(), tp.end() - 1);
    std::vector<Scalar> tv(values_.size());
    for (Index r = 0; r < num_rows_; ++r)
        for (Index k = row_ptr_[r]; k < row_ptr_[r + 1]; ++k) {
            const Index d = cur[col_idx_[k]]++;
            ti[d] = r;
            tv[d] = values_[k];
        }
    return CSRMatrix(std::move(tp), std::move(ti), std::move(tv), num_rows_);
}

namespace {
struct Fnv {
    std::uint64_t h = 14695981039346656037ULL;
    void word(std::uint64_t v) {
        h ^= v;
        h *= 1099511628211ULL;
    }
    void real(double v) {
        if (v == 0.0) v = 0.0;  // -0 hashes as +0
        std::uint64_t b;
        std::memcpy(&b, &v, sizeof b);
        word(b);
    }
};
}  // namespace

std::uint64_t computeStructuralFingerprint(const ModelIR& m) {
    Fnv f;
    f.word(static_cast<std::uint64_t>(m.numVars()));
    f.word(static_cast<std::uint64_t>(m.numRows()));
    f.word(static_cast<std::uint64_t>(m.obj_sense));
    for (double c : m.obj_coeffs) f.real(c);
    for (auto v : m.A.rowPtr()) f.word(static_cast<std::uint64_t>(v));
    for (auto v : m.A.colIdx()) f.word(static_cast<std::uint64_t>(v));
    for (double v : m.A.values()) f.real(v);
    for (CSRMatrix::Index j = 0; j < m.numVars(); ++j) {
        f.word(m.var_lower[j] <= -kInfinity ? 1 : 0);
        f.word(m.var_upper[j] >= kInfinity ? 1 : 0);
    }
    for (CSRMatrix::Index r = 0; r < m.numRows(); ++r) {
        f.word(m.isEqualityRow(r) ? 1 : 0);
        f.word(m.isFreeRow(r) ? 1 : 0);
        f.word(m.row_lower[r] <= -kInfinity ? 1 : 0);
        f.word(m.row_upper[r] >= kInfinity ? 1 : 0);
    }
    return f.h;
}

std::uint64_t computeModelFingerprint(const ModelIR& m) {
    Fnv f;
    f.word(static_cast<std::uint64_t>(m.numVars()));
    f.word(static_cast<std::uint64_t>(m.numRows()));
    f.real(m.obj_offset);
    f.word(static_cast<std::uint64_t>(m.obj_sense));
    for (double c : m.obj_coeffs) f.real(c);
    for (auto v : m.A.rowPtr()) f.word(static_cast<std::uint64_t>(v));
    for (auto v : m.A.colIdx()) f.word(static_cast<std::uint64_t>(v));
    for (double v : m.A.values()) f.real(v);
    for (CSRMatrix::Index j = 0; j < m.numVars(); ++j) {
        f.real(m.var_lower[j]);
        f.real(m.var_upper[j]);
        f.word(static_cast<std::uint64_t>(m.var_types[j]));
    }
    for (CSRMatrix::Index r = 0; r < m.numRows(); ++r) {
        f.real(m.row_lower[r]);
        f.real(m.row_upper[r]);
    }
    return f.h;
}

}  // namespace pramaan
