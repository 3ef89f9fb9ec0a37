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
#include <algorithm>
#include <cstring>
#include <vector>
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

// ------------------------------------------------------------------ batched forms
// The H stage and the MSM scalar preparation of a server step touch every
// request's vectors with the same few operations; each operation runs as one
// launch over a descriptor table (lane space = the concatenation of the
// requests' lanes, a lane finds its request by binary search on `begin`).

template <class D>
HZK_DEV uint32_t find_task(const D* __restrict__ t, uint32_t nt, uint64_t gid) {
    uint32_t lo = 0, hi = nt;   // t[lo].begin <= gid < t[hi].begin
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (t[mid].begin <= gid) lo = mid; else hi = mid;
    }
    return lo;
}

struct SpmvTask {
    CsrDev m[3];
    const fr_t* z;
    fr_t* out[3];
    uint64_t rows, begin;   // lanes: 3 * rows (matrix-major), rows = the largest row count
};

__global__ void __launch_bounds__(256) k_spmv3_batch(const SpmvTask* __restrict__ tasks, uint32_t nt, uint64_t total) {
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        const SpmvTask& t = tasks[find_task(tasks, nt, gid)];
        const uint64_t local = gid - t.begin, sel = local / t.rows, r = local - sel * t.rows;
        const CsrDev& m = t.m[sel];
        if (r >= m.row) continue;
        const uint64_t b = m.row_ptr[r], e = r + 1 < m.row ? m.row_ptr[r + 1] : m.cnt;
        fr_t acc = fr_t::zero();
        for (uint64_t j = b; j < e; ++j) acc += t.z[m.col[j]] * m.val[j];
        t.out[sel][r] = acc;
    }
}

struct AbcTask {
    fr_t* A;
    const fr_t* B;
    const fr_t* C;
    const fr_t* scale;
    uint64_t begin;
};

__global__ void __launch_bounds__(256) k_abc_batch(const AbcTask* __restrict__ tasks, uint32_t nt, uint64_t total) {
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        const AbcTask& t = tasks[find_task(tasks, nt, gid)];
        const uint64_t i = gid - t.begin;
        fr_t v = t.A[i] * t.B[i] - t.C[i];
        if (t.scale) v = v * *t.scale;
        t.A[i] = v;
    }
}

struct GatherTask {
    const uint4* src;
    const uint32_t* idx;   // null: contiguous
    uint4* dst;
    uint64_t begin;        // lanes: n * q words, or n elements when scaled
    fr_t r;
    uint32_t q, scaled;
};

__global__ void __launch_bounds__(256) k_gather_batch(const GatherTask* __restrict__ tasks, uint32_t nt, uint64_t total) {
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        const GatherTask& t = tasks[find_task(tasks, nt, gid)];
        const uint64_t local = gid - t.begin;
        if (t.scaled) {   // one Fr element per lane
            const uint64_t s = t.idx ? t.idx[local] : local;
            ((fr_t*)t.dst)[local] = ((const fr_t*)t.src)[s] * t.r;
        } else {          // one 16-byte word per lane: coalesced writes
            const uint64_t e = local / t.q, w = local - e * t.q;
            const uint64_t s = t.idx ? t.idx[e] : e;
            t.dst[local] = t.src[s * t.q + w];
        }
    }
}

template <class Task>
static int launch_tasks(hzk_ctx* ctx, const std::vector<Task>& tasks, uint64_t total, void (*kern)(const Task*, uint32_t, uint64_t),
                        const char* name) {
    if (tasks.empty() || total == 0) return HZK_OK;
    const void* dev = nullptr;
    int rc = upload_desc(ctx, tasks.data(), tasks.size() * sizeof(Task), &dev);
    if (rc) return rc;
    const uint64_t blocks = (total + 255) / 256, cap = (uint64_t)ctx->num_cus * 32;
    TimedLaunch tl(ctx, name);
    hipLaunchKernelGGL(kern, dim3((unsigned)(blocks < cap ? blocks : cap)), dim3(256), 0, ctx->stream, (const Task*)dev,
                       (uint32_t)tasks.size(), total);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

int spmv3_batch(hzk_ctx* ctx, const SpmvJob* jobs, size_t count) {
    std::vector<SpmvTask> tasks;
    uint64_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        HZK_CHECK(ctx, jobs[i].z && jobs[i].out[0] && jobs[i].out[1] && jobs[i].out[2], HZK_ERR_INVALID, "spmv3: null argument");
        SpmvTask t{};
        t.z = (const fr_t*)jobs[i].z;
        for (int k = 0; k < 3; ++k) {
            const hzk_csr* m = jobs[i].m[k];
            HZK_CHECK(ctx, m && (m->row == 0 || (m->row_ptr && (m->cnt == 0 || (m->val && m->col_idx)))), HZK_ERR_INVALID,
                      "spmv3: matrix %d has null arrays", k);
            t.m[k] = {(const fr_t*)m->val, m->col_idx, m->row_ptr, m->row, m->cnt};
            t.out[k] = (fr_t*)jobs[i].out[k];
            t.rows = std::max<uint64_t>(t.rows, m->row);
        }
        if (!t.rows) continue;
        t.begin = total;
        total += 3 * t.rows;
        tasks.push_back(t);
    }
    return launch_tasks(ctx, tasks, total, k_spmv3_batch, "spmv3");
}

int pointwise_abc_batch(hzk_ctx* ctx, const AbcJob* jobs, size_t count) {
    std::vector<AbcTask> tasks;
    uint64_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        HZK_CHECK(ctx, jobs[i].A && jobs[i].B && jobs[i].C, HZK_ERR_INVALID, "pointwise_abc: null argument");
        if (!jobs[i].n) continue;
        tasks.push_back({(fr_t*)jobs[i].A, (const fr_t*)jobs[i].B, (const fr_t*)jobs[i].C, (const fr_t*)jobs[i].scale, total});
        total += jobs[i].n;
    }
    return launch_tasks(ctx, tasks, total, k_abc_batch, "pointwise_abc");
}

int gather_batch(hzk_ctx* ctx, const GatherJob* jobs, size_t count) {
    std::vector<GatherTask> tasks;
    uint64_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        const GatherJob& j = jobs[i];
        HZK_CHECK(ctx, j.src && j.dst && j.elem_bytes % 16 == 0 && j.elem_bytes > 0 && (!j.r_host || j.elem_bytes == 32),
                  HZK_ERR_INVALID, "gather: bad argument");
        if (!j.n) continue;
        GatherTask t{};
        t.src = (const uint4*)j.src;
        t.idx = j.idx;
        t.dst = (uint4*)j.dst;
        t.q = (uint32_t)(j.elem_bytes / 16);
        t.scaled = j.r_host != nullptr;
        if (t.scaled) std::memcpy(t.r.l, j.r_host, 32);
        t.begin = total;
        total += t.scaled ? j.n : j.n * t.q;
        tasks.push_back(t);
    }
    return launch_tasks(ctx, tasks, total, k_gather_batch, "gather_rows");
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
