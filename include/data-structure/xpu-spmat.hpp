// CSR matrices with the interface of the reference's xpu::spmat
// (include/data-structure/xpu-spmat.cuh:5-105): spmat_descript (device
// arrays; row_ptr holds `row` entries and no sentinel, the last row ends at
// cnt), next_pow2, and spmat's MatMul generator for the reference's own
// statement (application_scence::MatMul, matrix_type A / B / C, parameter p).
//
// The MatMul statement proves Z = X Y for p x p matrices over the variable
// vector z = 1 | X (p^2) | Y (p^2) | Z (p^2) | T (p^3), T_{k,i,j} = X_{i,k} Y_{k,j}:
// p^3 product rows (A picks X, B picks Y, C picks T) then p^2 sum rows
// (A sums the p products of one Z entry, B picks the constant 1, C picks the
// Z entry), padded with empty rows to M = next_pow2(p^3 + p^2). The column
// formulas below are that statement's variable numbering, which proofs of the
// same statement must share.
#pragma once
#include <cstddef>
#include <cstdint>
#include "xpu-vector.hpp"

namespace xpu {

inline size_t next_pow2(size_t x) {
    size_t r = 1;
    while (r < x) r <<= 1;
    return r;
}

enum class application_scence { MatMul };
enum class matrix_type { A, B, C };

template <typename FieldT>
struct spmat_descript {
    FieldT* val;
    size_t *col_idx, *row_ptr;
    size_t row, col, cnt;
};

template <typename FieldT>
class spmat {
    vector<FieldT> val_;
    vector<size_t> col_idx_, row_ptr_;
    size_t row_ = 0, col_ = 0, cnt_ = 0;

    void matmul(matrix_type mt, size_t p) {
        const size_t p2 = p * p, p3 = p2 * p, M = next_pow2(p3 + p2);
        const size_t nnz = mt == matrix_type::A ? 2 * p3 : p3 + p2;
        const size_t x0 = 1, y0 = 1 + p2, z0 = 1 + 2 * p2, t0 = 1 + 3 * p2;
        val_.allocate(nnz, mem_policy::cross_platform);
        col_idx_.allocate(nnz, mem_policy::cross_platform);
        row_ptr_.allocate(M, mem_policy::cross_platform);
        row_ = 0, col_ = t0 + p3, cnt_ = 0;
        auto put = [&](size_t c) { val_[cnt_] = FieldT::one(), col_idx_[cnt_++] = c; };
        // product rows: t = (k, i, j) with t = k p^2 + i p + j
        for (size_t t = 0; t < p3; ++t) {
            const size_t k = t / p2, i = (t / p) % p, j = t % p;
            row_ptr_[row_++] = cnt_;
            put(mt == matrix_type::A ? x0 + i * p + k : mt == matrix_type::B ? y0 + k * p + j : t0 + t);
        }
        // sum rows: Z_e = sum_k T_{k, e}, e = i p + j
        for (size_t e = 0; e < p2; ++e) {
            row_ptr_[row_++] = cnt_;
            if (mt == matrix_type::A)
                for (size_t k = 0; k < p; ++k) put(t0 + k * p2 + e);
            else
                put(mt == matrix_type::B ? 0 : z0 + e);
        }
        while (row_ < M) row_ptr_[row_++] = cnt_;
        val_.store(), col_idx_.store(), row_ptr_.store();
    }

public:
    spmat() = default;
    spmat(application_scence as, matrix_type mt, size_t param1) { init(as, mt, param1); }
    void init(application_scence as, matrix_type mt, size_t param1) {
        if (as == application_scence::MatMul) matmul(mt, param1);
    }
    template <typename devFdT>
    spmat_descript<devFdT> descript() {
        return {(devFdT*)val_.p(), (size_t*)col_idx_.p(), (size_t*)row_ptr_.p(), row_, col_, cnt_};
    }
    size_t rows() const { return row_; }
    size_t cols() const { return col_; }
    size_t nnz() const { return cnt_; }
    const vector<size_t>& host_col_idx() const { return col_idx_; }
    const vector<size_t>& host_row_ptr() const { return row_ptr_; }
};

}  // namespace xpu
