// R1CS sparse matrix-vector products and pointwise polynomial arithmetic.
//
// Replaces batch_sparse_matrix_vector_multiplication
// (include/operator/poly.cuh:5-30) and polynomial_A_times_B_minus_C
// (poly.cuh:75-86). The CSR layout is the reference's spmat_descript
// (include/data-structure/xpu-spmat.cuh:20-25): val Fr[cnt], col_idx and
// row_ptr as 64-bit indices, row_ptr holding `row` entries with no sentinel
// (the last row ends at cnt).
//
// One lane per row over the three matrices in one launch (synthetic R1CS rows
// hold 1-3 nonzeros, so a row is a handful of gathers and Fr products; SpMV is
// gather-bound, ~68 B per nonzero).
#include <cstring>
#include "common.hpp"
#include "field.hpp"

namespace hzk {

struct CsrDev {
    const fr_t* val;
    const uint64_t* col;
    const uint64_t* row_ptr;
    uint64_t row, cnt;
};

__global__ void __launch_bounds__(256) k_spmv3(CsrDev A, CsrDev B, CsrDev C, const fr_t* __restrict__ z, fr_t* __restrict__ Az,
                                               fr_t* __restrict__ Bz, fr_t* __restrict__ Cz) {
    const uint32_t sel = blockIdx.y;
    const CsrDev& m = sel == 0 ? A : (sel == 1 ? B : C);
    fr_t* out = sel == 0 ? Az : (sel == 1 ? Bz : Cz);
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m.row) return;
    size_t b = m.row_ptr[r], e = r + 1 < m.row ? m.row_ptr[r + 1] : m.cnt;
    fr_t acc = fr_t::zero();
    for (size_t j = b; j < e; ++j) acc += z[m.col[j]] * m.val[j];
    out[r] = acc;
}

// A <- A o B - C (o: pointwise product), optionally times a scalar
__global__ void __launch_bounds__(256) k_abc(fr_t* __restrict__ A, const fr_t* __restrict__ B, const fr_t* __restrict__ C,
                                             size_t n, const fr_t* __restrict__ scale) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t v = A[i] * B[i] - C[i];
    if (scale) v = v * *scale;
    A[i] = v;
}

int spmv3(hzk_ctx* ctx, const hzk_csr* A, const hzk_csr* B, const hzk_csr* C, const void* z, void* Az, void* Bz, void* Cz) {
    HZK_CHECK(ctx, A && B && C && z && Az && Bz && Cz, HZK_ERR_INVALID, "spmv3: null argument");
    const hzk_csr* ms[3] = {A, B, C};
    CsrDev d[3];
    uint64_t rows = 0;
    for (int i = 0; i < 3; ++i) {
        HZK_CHECK(ctx, ms[i]->row == 0 || (ms[i]->row_ptr && (ms[i]->cnt == 0 || (ms[i]->val && ms[i]->col_idx))),
                  HZK_ERR_INVALID, "spmv3: matrix %d has null arrays", i);
        d[i] = {(const fr_t*)ms[i]->val, ms[i]->col_idx, ms[i]->row_ptr, ms[i]->row, ms[i]->cnt};
        rows = rows > ms[i]->row ? rows : ms[i]->row;
    }
    if (rows == 0) return HZK_OK;
    TimedLaunch tl(ctx, "spmv3");
    hipLaunchKernelGGL(k_spmv3, dim3(blocks_for(rows, 256), 3), dim3(256), 0, ctx->stream, d[0], d[1], d[2], (const fr_t*)z,
                       (fr_t*)Az, (fr_t*)Bz, (fr_t*)Cz);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

// dst[k] = src[idx[k]] for elements of q 16-byte words: thread per word, so
// the writes are coalesced and each element's words are read together
__global__ void __launch_bounds__(256) k_gather_rows(const uint4* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                                                     uint32_t q, uint4* __restrict__ dst) {
    const uint64_t total = n * q;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = g / q;
        dst[g] = src[(uint64_t)idx[e] * q + (g - e * q)];
    }
}

int gather_rows(hzk_ctx* ctx, const void* src, const uint32_t* idx, size_t n, size_t elem_bytes, void* dst) {
    HZK_CHECK(ctx, src && idx && dst && elem_bytes % 16 == 0 && elem_bytes > 0, HZK_ERR_INVALID, "gather_rows: bad argument");
    if (n == 0) return HZK_OK;
    const uint32_t q = (uint32_t)(elem_bytes / 16);
    const uint64_t words = (uint64_t)n * q;
    const uint64_t blocks = (words + 255) / 256;
    const unsigned grid = (unsigned)(blocks < (uint64_t)ctx->num_cus * 32 ? blocks : (uint64_t)ctx->num_cus * 32);
    TimedLaunch tl(ctx, "gather_rows");
    hipLaunchKernelGGL(k_gather_rows, dim3(grid), dim3(256), 0, ctx->stream, (const uint4*)src, idx, (uint64_t)n, q, (uint4*)dst);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

// dst[k] = z[idx[k]] * r (Montgomery product)
__global__ void __launch_bounds__(256) k_gather_scale(const fr_t* __restrict__ z, const uint32_t* __restrict__ idx, size_t n,
                                                      fr_t r, fr_t* __restrict__ dst) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = z[idx[k]] * r;
}

int gather_scale_fr(hzk_ctx* ctx, const void* z, const uint32_t* idx, size_t n, const void* r_host, void* dst) {
    HZK_CHECK(ctx, z && idx && r_host && dst, HZK_ERR_INVALID, "gather_scale_fr: null argument");
    if (n == 0) return HZK_OK;
    fr_t r;
    std::memcpy(r.l, r_host, 32);
    const uint64_t blocks = (n + 255) / 256;
    const unsigned grid = (unsigned)(blocks < (uint64_t)ctx->num_cus * 16 ? blocks : (uint64_t)ctx->num_cus * 16);
    TimedLaunch tl(ctx, "gather_rows");
    hipLaunchKernelGGL(k_gather_scale, dim3(grid), dim3(256), 0, ctx->stream, (const fr_t*)z, idx, n, r, (fr_t*)dst);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

int pointwise_abc(hzk_ctx* ctx, size_t n, void* A, const void* B, const void* C, const void* scale) {
    HZK_CHECK(ctx, A && B && C, HZK_ERR_INVALID, "pointwise_abc: null argument");
    if (n == 0) return HZK_OK;
    TimedLaunch tl(ctx, "pointwise_abc");
    hipLaunchKernelGGL(k_abc, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, (fr_t*)A, (const fr_t*)B, (const fr_t*)C, n,
                       (const fr_t*)scale);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

}  // namespace hzk

using namespace hzk;

extern "C" {

int hzk_spmv3(hzk_ctx* ctx, const hzk_csr* A, const hzk_csr* B, const hzk_csr* C, const void* z, void* Az, void* Bz, void* Cz) {
    if (!ctx) return HZK_ERR_INVALID;
    return spmv3(ctx, A, B, C, z, Az, Bz, Cz);
}

int hzk_pointwise_abc(hzk_ctx* ctx, size_t deg, void* A, const void* B, const void* C) {
    if (!ctx) return HZK_ERR_INVALID;
    return pointwise_abc(ctx, deg, A, B, C, nullptr);
}

}  // extern "C"
