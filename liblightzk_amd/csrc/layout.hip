// Base-table layouts of the reference and conversion to this library's own.
//
// The reference keeps MSM bases as level-major fixed-base tables built by
// pre_comp_g1/_g2 (include/operator/msm.cuh:48-75): record i + j*length holds
// 2^(gap*j) * P_i as a 96-byte Jacobian record X|Y|Z normalised to Z = 1;
// for G2 every point is two 96-byte lane records, [X.c0|Y.c0|Z.c0] then
// [X.c1|Y.c1|Z.c1] (g2_t::read_from/write_to,
// include/device-arith/curve/alt_bn128_t.cuh:244-256), so level j starts at
// record 2*j*length. The point at infinity leaves to_affine as (0, 0, 1)
// (jacobian_t.cuh:36-57 with reciprocal(0) = 0), whose x and y are exactly
// this library's affine infinity (0, 0).
//
// Here an MSM needs only level 0 (Sum s_i P_i, computed window by window), so
// a reference table is read by converting its first `length` points into the
// 64 / 128-byte affine layout. hzk_precomp rebuilds the reference's table in
// place (level 0 normalised, levels 1.. by doubling), for callers that still
// hand reference-layout tables around. hzk_bases_levels builds this library's
// own level table (affine, level-major) for the single-bucket-set MSM.
#include "common.hpp"
#include "curve.hpp"

namespace hzk {

// one coordinate c (0 X, 1 Y, 2 Z) of point i in a reference record table
HZK_DEV fq_t ref_coord(const fq_t* rec, size_t i, int c, fq_t*) { return rec[3 * i + c]; }
HZK_DEV fq2_t ref_coord(const fq_t* rec, size_t i, int c, fq2_t*) { return {rec[6 * i + c], rec[6 * i + 3 + c]}; }
HZK_DEV void ref_store(fq_t* rec, size_t i, const fq_t& x, const fq_t& y, const fq_t& z) {
    rec[3 * i] = x;
    rec[3 * i + 1] = y;
    rec[3 * i + 2] = z;
}
HZK_DEV void ref_store(fq_t* rec, size_t i, const fq2_t& x, const fq2_t& y, const fq2_t& z) {
    rec[6 * i] = x.c0;
    rec[6 * i + 1] = y.c0;
    rec[6 * i + 2] = z.c0;
    rec[6 * i + 3] = x.c1;
    rec[6 * i + 4] = y.c1;
    rec[6 * i + 5] = z.c1;
}

template <class F>
__global__ void __launch_bounds__(256) k_ref_to_affine(const fq_t* __restrict__ rec, affine_t<F>* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = {ref_coord(rec, i, 0, (F*)nullptr), ref_coord(rec, i, 1, (F*)nullptr)};
}

template <class F>
__global__ void __launch_bounds__(256) k_affine_to_ref(const affine_t<F>* __restrict__ in, fq_t* __restrict__ rec, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    affine_t<F> a = in[i];
    ref_store(rec, i, a.x, a.y, F::one());
}

// XYZZ -> affine with one inversion; infinity -> (0, 0)
template <class F>
HZK_DEV affine_t<F> to_affine_one(const xyzz_t<F>& p) {
    if (p.is_inf()) return affine_t<F>::inf();
    F t = (p.ZZ * p.ZZZ).inverse();
    return {p.X * (t * p.ZZZ), p.Y * (t * p.ZZ)};
}

// pre_comp semantics: level 0 normalised in place (Jacobian X/Z^2, Y/Z^3 with
// Z = 0 for infinity), record j*length + i = 2^(gap*j) P_i, each with Z = 1
template <class F>
__global__ void __launch_bounds__(128) k_precomp(fq_t* __restrict__ rec, size_t length, uint32_t gap, uint32_t level) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= length) return;
    F X = ref_coord(rec, i, 0, (F*)nullptr), Y = ref_coord(rec, i, 1, (F*)nullptr), Z = ref_coord(rec, i, 2, (F*)nullptr);
    xyzz_t<F> cur = xyzz_t<F>::inf();
    if (!Z.is_zero()) {
        F zz = Z.sqr();
        cur = {X, Y, zz, zz * Z};
    }
    for (uint32_t j = 0; j < level; ++j) {
        if (j)
            for (uint32_t k = 0; k < gap; ++k) cur = cur.dbl();
        affine_t<F> a = to_affine_one(cur);
        ref_store(rec, i + (size_t)j * length, a.x, a.y, F::one());
    }
}

// fixed-base level table: out[j*n + i] = 2^(c j) in[i], affine, j < L
template <class F>
__global__ void __launch_bounds__(128) k_levels(const affine_t<F>* __restrict__ in, affine_t<F>* __restrict__ out, size_t n,
                                                uint32_t c, uint32_t L) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const affine_t<F> a = in[i];
    out[i] = a;
    xyzz_t<F> cur = xyzz_t<F>::from_affine(a);
    for (uint32_t j = 1; j < L; ++j) {
        for (uint32_t k = 0; k < c; ++k) cur = cur.dbl();
        out[(size_t)j * n + i] = to_affine_one(cur);
    }
}

int convert_bases(hzk_ctx* ctx, int group, uint32_t from, uint32_t to, const void* src, void* dst, size_t n) {
    HZK_CHECK(ctx, group == 1 || group == 2, HZK_ERR_INVALID, "bases_convert: group %d", group);
    HZK_CHECK(ctx, src && dst && n > 0, HZK_ERR_INVALID, "bases_convert: null buffer or n = 0");
    HZK_CHECK(ctx, from <= HZK_BASES_REF_LEVELS && to <= HZK_BASES_REF_LEVELS && from != to, HZK_ERR_INVALID,
              "bases_convert: layouts %u -> %u", from, to);
    dim3 grid(blocks_for(n, 256)), block(256);
    if (from == HZK_BASES_REF_LEVELS) {
        if (group == 1) hipLaunchKernelGGL(k_ref_to_affine<fq_t>, grid, block, 0, ctx->stream, (const fq_t*)src, (affine_t<fq_t>*)dst, n);
        else hipLaunchKernelGGL(k_ref_to_affine<fq2_t>, grid, block, 0, ctx->stream, (const fq_t*)src, (affine_t<fq2_t>*)dst, n);
    } else {
        if (group == 1) hipLaunchKernelGGL(k_affine_to_ref<fq_t>, grid, block, 0, ctx->stream, (const affine_t<fq_t>*)src, (fq_t*)dst, n);
        else hipLaunchKernelGGL(k_affine_to_ref<fq2_t>, grid, block, 0, ctx->stream, (const affine_t<fq2_t>*)src, (fq_t*)dst, n);
    }
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

// level 0 of a reference table as affine bases in the context's conversion buffer
int ref_bases_affine(hzk_ctx* ctx, int group, const void* ref, size_t n, const void** out) {
    size_t bytes = n * (group == 1 ? 64 : 128);
    if (ctx->conv_bytes < bytes) {
        if (ctx->conv) {
            HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
            HZK_HIP(ctx, hipFree(ctx->conv));
            ctx->conv = nullptr;
            ctx->conv_bytes = 0;
        }
        hipError_t e = hipMalloc(&ctx->conv, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(ctx, HZK_ERR_NOMEM, "conversion buffer of %zu bytes: %s", bytes, hipGetErrorString(e));
        }
        ctx->conv_bytes = bytes;
    }
    int rc = convert_bases(ctx, group, HZK_BASES_REF_LEVELS, HZK_BASES_AFFINE, ref, ctx->conv, n);
    if (rc) return rc;
    *out = ctx->conv;
    return 0;
}

}  // namespace hzk

using namespace hzk;

extern "C" {

int hzk_bases_convert(hzk_ctx* ctx, int group, uint32_t from_layout, uint32_t to_layout, const void* src, void* dst, size_t n) {
    if (!ctx) return HZK_ERR_INVALID;
    return convert_bases(ctx, group, from_layout, to_layout, src, dst, n);
}

int hzk_bases_levels(hzk_ctx* ctx, int group, int c, int levels, const void* bases, size_t n, void* out) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, group == 1 || group == 2, HZK_ERR_INVALID, "bases_levels: group %d", group);
    HZK_CHECK(ctx, c >= 4 && c <= 16 && bases && out && n > 0, HZK_ERR_INVALID, "bases_levels: c %d, n %zu", c, n);
    HZK_CHECK(ctx, levels >= 1 && levels <= (255 + c - 1) / c, HZK_ERR_INVALID, "bases_levels: %d levels for c %d", levels, c);
    const uint32_t L = (uint32_t)levels;
    dim3 grid(blocks_for(n, 128)), block(128);
    if (group == 1)
        hipLaunchKernelGGL(k_levels<fq_t>, grid, block, 0, ctx->stream, (const affine_t<fq_t>*)bases, (affine_t<fq_t>*)out, n,
                           (uint32_t)c, L);
    else
        hipLaunchKernelGGL(k_levels<fq2_t>, grid, block, 0, ctx->stream, (const affine_t<fq2_t>*)bases, (affine_t<fq2_t>*)out, n,
                           (uint32_t)c, L);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

int hzk_precomp(hzk_ctx* ctx, int group, void* points, size_t length, size_t gap, size_t level) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, group == 1 || group == 2, HZK_ERR_INVALID, "precomp: group %d", group);
    HZK_CHECK(ctx, points && length > 0 && level >= 1 && level <= 256 && gap <= 64, HZK_ERR_INVALID,
              "precomp: length %zu, gap %zu, level %zu", length, gap, level);
    dim3 grid(blocks_for(length, 128)), block(128);
    if (group == 1) hipLaunchKernelGGL(k_precomp<fq_t>, grid, block, 0, ctx->stream, (fq_t*)points, length, (uint32_t)gap, (uint32_t)level);
    else hipLaunchKernelGGL(k_precomp<fq2_t>, grid, block, 0, ctx->stream, (fq_t*)points, length, (uint32_t)gap, (uint32_t)level);
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

}  // extern "C"
