// Drop-in for the reference's include/protocol/groth16-operator.cuh: the
// twiddle-table builder ntt_setup and the operator headers it pulls in.
#pragma once
#include <cstddef>
#include "../alt_bn128_types.hpp"
#include "../data-structure/xpu-spmat.hpp"
#include "../data-structure/xpu-vector.hpp"
#include "../lightzk_context.hpp"
#include "../operator/msm.hpp"
#include "../operator/poly.hpp"

// include/protocol/groth16-operator.cuh:6-32: ntt_params <- 6M + 2 Fr
// elements (mirrored on host and device): [0, M) w_M^i, [M, 2M) w_M^-i,
// 2M: M^-1, [2M+1, 4M+1) w_2M^i, [4M+1, 6M+1) w_2M^-i and, at 6M + 1, the
// value the reference writes there, (4M)^-1. The roots are the library's
// (the primitive 2^k-th roots 5^((r-1)/2^k) of the reference's tables); the
// config's string tables are not read.
template <typename FieldT, ntt_config* nc>
void ntt_setup(size_t M, xpu::vector<FieldT>& ntt_params) {
    static_assert(sizeof(FieldT) == 32, "ntt_setup: 32-byte Fr elements");
    (void)nc;
    ntt_params.allocate(6 * M + 2, xpu::mem_policy::cross_platform);
    int rc = hzk_ntt_setup(M, &ntt_params[0]);
    if (rc) throw lightzk::error(rc, "ntt_setup: M must be a power of two <= 2^26");
    ntt_params.store();
}
