// Drop-in for the reference's include/operator/poly.cuh (CSR descriptors from
// include/data-structure/xpu-spmat.hpp).
#pragma once
#include <cstddef>
#include "../data-structure/xpu-spmat.hpp"
#include "../lightzk_context.hpp"

namespace lightzk {
template <typename T>
inline hzk_csr to_csr(const xpu::spmat_descript<T>& d) {
    static_assert(sizeof(size_t) == sizeof(uint64_t), "64-bit indices");
    return hzk_csr{d.val, (const uint64_t*)d.col_idx, (const uint64_t*)d.row_ptr, d.row, d.col, d.cnt};
}
}  // namespace lightzk

// include/operator/poly.cuh:5-30 (Az, Bz, Cz need not be zeroed here)
template <typename devFdT>
void batch_sparse_matrix_vector_multiplication(xpu::spmat_descript<devFdT> A, xpu::spmat_descript<devFdT> B,
                                               xpu::spmat_descript<devFdT> C, devFdT* z, devFdT* Az, devFdT* Bz,
                                               devFdT* Cz) {
    hzk_csr a = lightzk::to_csr(A), b = lightzk::to_csr(B), c3 = lightzk::to_csr(C);
    lightzk::run([&](hzk_ctx* c) { return hzk_spmv3(c, &a, &b, &c3, z, Az, Bz, Cz); });
}

// include/operator/poly.cuh:65-73: coeffs[k] <- factor * sum_j coeffs[j] w^(jk),
// omega_powers[e] = w^e for e < deg/2, factor nullable
template <typename devFdT>
void number_theory_transformation(size_t deg, size_t log_deg, devFdT* coeffs, devFdT* omega_powers, devFdT* factor) {
    if (((size_t)1 << log_deg) != deg) throw lightzk::error(HZK_ERR_INVALID, "number_theory_transformation: deg != 2^log_deg");
    lightzk::run([&](hzk_ctx* c) { return hzk_ntt(c, coeffs, deg, omega_powers, 0, factor); });
}

// include/operator/poly.cuh:75-86
template <typename devFdT>
void polynomial_A_times_B_minus_C(size_t deg, devFdT* A, devFdT* B, devFdT* C) {
    lightzk::run([&](hzk_ctx* c) { return hzk_pointwise_abc(c, deg, A, B, C); });
}
