// sparse_matrix.hpp -- CSR matrix of the PRAMAAN model interface.
//
// Same public interface as the reference's CSRMatrix
// (include/pramaan/sparse_matrix.hpp:28-181): int32 indices, FP64 values,
// column indices sorted and unique within a row, O(nnz) multiply, row
// views, column lookups and a counting-sort transpose.  Code written
// against the reference header compiles against this one unchanged.
//
// multiply() sums each row left to right starting from 0.0, the order the
// certificate writer's row activities depend on for byte-identical files.
#pragma once
#include <algorithm>
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace pramaan {

class CSRMatrix {
public:
    using Index = std::int32_t;
    using Scalar = double;

    CSRMatrix() = default;
    CSRMatrix(std::vector<Index> row_ptr, std::vector<Index> col_idx, std::vector<Scalar> values, Index num_cols)
        : row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)), values_(std::move(values)), num_cols_(num_cols) {
        num_rows_ = row_ptr_.empty() ? 0 : static_cast<Index>(row_ptr_.size() - 1);
        require_consistent();
    }

    Index numRows() const noexcept { return num_rows_; }
    Index numCols() const noexcept { return num_cols_; }
    Index nnz() const noexcept { return static_cast<Index>(values_.size()); }
    const std::vector<Index>& rowPtr() const noexcept { return row_ptr_; }
    const std::vector<Index>& colIdx() const noexcept { return col_idx_; }
    const std::vector<Scalar>& values() const noexcept { return values_; }

    std::vector<Scalar> multiply(const std::vector<Scalar>& x) const {
        std::vector<Scalar> y;
        multiplyInto(x, y);
        return y;
    }
    // y = A x; y is resized to numRows().
    void multiplyInto(const std::vector<Scalar>& x, std::vector<Scalar>& y) const {
        if (x.size() != static_cast<std::size_t>(num_cols_))
            throw std::invalid_argument("CSRMatrix::multiply: x has " + std::to_string(x.size()) +
                                        " entries, the matrix " + std::to_string(num_cols_) + " columns");
        y.resize(static_cast<std::size_t>(num_rows_));
        const Index* col = col_idx_.data() + (num_rows_ > 0 ? row_ptr_[0] : 0);
        const Scalar* val = values_.data() + (num_rows_ > 0 ? row_ptr_[0] : 0);
        for (Index r = 0; r < num_rows_; ++r) {
            const Index* const stop = col_idx_.data() + row_ptr_[r + 1];
            Scalar acc = 0.0;
            while (col != stop) acc += *val++ * x[*col++];
            y[r] = acc;
        }
    }

    struct Entry {
        Index index;
        Scalar value;
    };

    // The entries of one row, in increasing column order: a pair of
    // parallel pointers that advance together.
    class RowView {
    public:
        RowView(const Index* cols, const Scalar* vals, Index count) : col_(cols), val_(vals), count_(count) {}
        struct Iterator {
            const Index* col;
            const Scalar* val;
            Entry operator*() const { return {*col, *val}; }
            Iterator& operator++() {
                ++col;
                ++val;
                return *this;
            }
            bool operator!=(const Iterator& other) const { return col != other.col; }
        };
        Iterator begin() const { return {col_, val_}; }
        Iterator end() const { return {col_ + count_, val_ + count_}; }
        Index size() const noexcept { return count_; }

    private:
        const Index* col_;
        const Scalar* val_;
        Index count_;
    };

    RowView row(Index r) const {
        assert(0 <= r && r < num_rows_);
        const Index b = row_ptr_[r];
        return {col_idx_.data() + b, values_.data() + b, row_ptr_[r + 1] - b};
    }

    // Entries of column c as (row, value), rows increasing: one binary
    // search per row over its sorted column indices.
    std::vector<Entry> column(Index c) const {
        assert(0 <= c && c < num_cols_);
        std::vector<Entry> hits;
        for (Index r = 0; r < num_rows_; ++r) {
            const Index* lo = col_idx_.data() + row_ptr_[r];
            const Index* hi = col_idx_.data() + row_ptr_[r + 1];
            const Index* at = std::lower_bound(lo, hi, c);
            if (at != hi && *at == c) hits.push_back({r, values_[static_cast<std::size_t>(at - col_idx_.data())]});
        }
        return hits;
    }

    CSRMatrix transpose() const;

private:
    // Rejects a CSR whose arrays disagree with each other or with the
    // column count.
    void require_consistent() const {
        const auto fail = [](const std::string& why) { throw std::invalid_argument("CSRMatrix: " + why); };
        const std::size_t nnz = col_idx_.size();
        if (row_ptr_.size() != static_cast<std::size_t>(num_rows_) + 1)
            fail("expected " + std::to_string(num_rows_ + 1) + " row pointers, got " + std::to_string(row_ptr_.size()));
        if (values_.size() != nnz)
            fail(std::to_string(nnz) + " column indices but " + std::to_string(values_.size()) + " values");
        if (!row_ptr_.empty() && static_cast<std::size_t>(row_ptr_.back()) != nnz)
            fail("the last row pointer must equal the number of entries (" + std::to_string(nnz) + ")");
        if (std::adjacent_find(row_ptr_.begin(), row_ptr_.end(), [](Index a, Index b) { return b < a; }) !=
            row_ptr_.end())
            fail("a row pointer decreases");
        const auto bad = std::find_if(col_idx_.begin(), col_idx_.end(),
                                      [this](Index c) { return c < 0 || c >= num_cols_; });
        if (bad != col_idx_.end())
            fail("column index " + std::to_string(*bad) + " outside [0, " + std::to_string(num_cols_) + ")");
    }

    std::vector<Index> row_ptr_;
    std::vector<Index> col_idx_;
    std::vector<Scalar> values_;
    Index num_rows_ = 0;
    Index num_cols_ = 0;
};

}  // namespace pramaan
