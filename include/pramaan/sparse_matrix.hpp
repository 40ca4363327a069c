// sparse_matrix.hpp -- CSR matrix of the PRAMAAN model interface.
//
// Same public interface as the reference's CSRMatrix
// (include/pramaan/sparse_matrix.hpp:28-181): int32 indices, FP64 values,
// column indices sorted and unique within a row, O(nnz) multiply, row
// views, column scans and a counting-sort transpose.  Code written against
// the reference header compiles against this one unchanged.
#pragma once
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace pramaan {

class CSRMatrix {
public:
    using Index = std::int32_t;
    using Scalar = double;

    CSRMatrix() = default;
    CSRMatrix(std::vector<Index> row_ptr, std::vector<Index> col_idx, std::vector<Scalar> values, Index num_cols)
        : row_ptr_(std::move(row_ptr)), col_idx_(std::move(col_idx)), values_(std::move(values)), num_cols_(num_cols) {
        num_rows_ = row_ptr_.empty() ? 0 : static_cast<Index>(row_ptr_.size()) - 1;
        check();
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
    void multiplyInto(const std::vector<Scalar>& x, std::vector<Scalar>& y) const {
        if (static_cast<Index>(x.size()) != num_cols_)
            throw std::invalid_argument("CSRMatrix::multiply: x size != num_cols");
        y.assign(static_cast<std::size_t>(num_rows_), Scalar{0});
        for (Index r = 0; r < num_rows_; ++r) {
            Scalar s = 0;
            for (Index k = row_ptr_[r]; k < row_ptr_[r + 1]; ++k) s += values_[k] * x[col_idx_[k]];
            y[r] = s;
        }
    }

    struct Entry {
        Index index;
        Scalar value;
    };

    class RowView {
    public:
        RowView(const Index* cols, const Scalar* vals, Index count) : cols_(cols), vals_(vals), count_(count) {}
        struct Iterator {
            const Index* cols;
            const Scalar* vals;
            Index pos;
            Entry operator*() const { return Entry{cols[pos], vals[pos]}; }
            Iterator& operator++() {
                ++pos;
                return *this;
            }
            bool operator!=(const Iterator& o) const { return pos != o.pos; }
        };
        Iterator begin() const { return Iterator{cols_, vals_, 0}; }
        Iterator end() const { return Iterator{cols_, vals_, count_}; }
        Index size() const noexcept { return count_; }

    private:
        const Index* cols_;
        const Scalar* vals_;
        Index count_;
    };

    RowView row(Index r) const {
        assert(r >= 0 && r < num_rows_);
        return RowView(col_idx_.data() + row_ptr_[r], values_.data() + row_ptr_[r], row_ptr_[r + 1] - row_ptr_[r]);
    }

    // O(nnz) scan; fine for occasional use.
    std::vector<Entry> column(Index c) const {
        std::vector<Entry> out;
        for (Index r = 0; r < num_rows_; ++r)
            for (Index k = row_ptr_[r]; k < row_ptr_[r + 1]; ++k)
                if (col_idx_[k] == c) {
                    out.push_back(Entry{r, values_[k]});
                    break;
                }
        return out;
    }

    CSRMatrix transpose() const;

private:
    void check() const {
        if (static_cast<Index>(row_ptr_.size()) != num_rows_ + 1)
            throw std::invalid_argument("CSRMatrix: row_ptr must have num_rows + 1 entries");
        if (col_idx_.size() != values_.size()) throw std::invalid_argument("CSRMatrix: col_idx and values size mismatch");
        if (!row_ptr_.empty() && static_cast<std::size_t>(row_ptr_.back()) != col_idx_.size())
            throw std::invalid_argument("CSRMatrix: row_ptr.back() must equal nnz");
        for (std::size_t r = 0; r + 1 < row_ptr_.size(); ++r)
            if (row_ptr_[r] > row_ptr_[r + 1]) throw std::invalid_argument("CSRMatrix: row_ptr must be non-decreasing");
        for (Index c : col_idx_)
            if (c < 0 || c >= num_cols_) throw std::invalid_argument("CSRMatrix: col_idx entry out of bounds");
    }

    std::vector<Index> row_ptr_;
    std::vector<Index> col_idx_;
    std::vector<Scalar> values_;
    Index num_rows_ = 0;
    Index num_cols_ = 0;
};

}  // namespace pramaan
