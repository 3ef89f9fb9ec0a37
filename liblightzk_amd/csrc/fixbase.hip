// Fixed-base scalar multiplication P_i = k_i * G for G1 and G2, and batch
// normalisation of XYZZ points to affine.
//
// Replaces fix_base_multi_scalar_multiplication_g1/_g2
// (include/operator/msm.cuh:5-46: per-thread MSB-first double-and-add, ~254
// doublings + ~127 additions per point, scalars converted in place, and a G2
// variant that adds the uninitialised output instead of the generator at
// msm.cuh:42) and the affine normalisation half of pre_comp_g1/_g2
// (msm.cuh:48-75). Here:
//   * a comb table T[j][d] = d * 2^(8j) * G (32 x 256 affine points) is built
//     once per context on the host and kept in HBM;
//   * a lane per point adds the 32 table entries selected by the bytes of its
//     canonical scalar (mixed additions into XYZZ), leaving the scalars
//     untouched;
//   * a lane per run of 32 points normalises with Montgomery's trick (one
//     Fermat inversion per run).
#include <vector>
#include "common.hpp"
#include "curve.hpp"
#include "host_ec.hpp"

namespace hzk {

static constexpr int COMB_WINDOWS = 32;
static constexpr int NORM_RUN = 32;

template <class F>
__global__ void __launch_bounds__(256) k_fixbase(const fr_t* __restrict__ scalars, size_t n,
                                                 const affine_t<F>* __restrict__ table, xyzz_t<F>* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t k = scalars[i].from_mont();
    xyzz_t<F> acc = xyzz_t<F>::inf();
#pragma unroll 4
    for (int j = 0; j < COMB_WINDOWS; ++j) {
        uint32_t d = (k.l[j >> 2] >> (8 * (j & 3))) & 0xffu;
        if (d) acc.add_affine(table[j * 256 + d]);
    }
    out[i] = acc;
}

// affine[i] = normalise(pts[i]); infinity becomes (0, 0). pre: scratch, n elements.
template <class F>
__global__ void __launch_bounds__(256) k_batch_affine(const xyzz_t<F>* __restrict__ pts, size_t n, F* __restrict__ pre,
                                                      affine_t<F>* __restrict__ out) {
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t b = r * NORM_RUN;
    if (b >= n) return;
    size_t e = b + NORM_RUN < n ? b + NORM_RUN : n;
    F acc = F::one();
    for (size_t i = b; i < e; ++i) {
        pre[i] = acc;
        const xyzz_t<F>& p = pts[i];
        if (!p.is_inf()) acc = acc * (p.ZZ * p.ZZZ);
    }
    F inv = acc.inverse();
    for (size_t i = e; i-- > b;) {
        xyzz_t<F> p = pts[i];
        if (p.is_inf()) { out[i] = affine_t<F>::inf(); continue; }
        F d = p.ZZ * p.ZZZ;
        F di = inv * pre[i];
        inv = inv * d;
        out[i] = {p.X * (di * p.ZZZ), p.Y * (di * p.ZZ)};
    }
}

template <class HostPt>
static std::vector<char> comb_table(const HostPt& gen) {
    typedef decltype(gen.X) HF;
    std::vector<HostPt> pts(COMB_WINDOWS * 256);
    HostPt base = gen;
    for (int j = 0; j < COMB_WINDOWS; ++j) {
        HostPt acc = HostPt::inf();
        pts[j * 256] = acc;
        for (int d = 1; d < 256; ++d) {
            acc = acc.add(base);
            pts[j * 256 + d] = acc;
        }
        for (int b = 0; b < 8; ++b) base = base.dbl();
    }
    std::vector<HF> xs, ys;
    host::batch_affine(pts, xs, ys);
    std::vector<char> bytes(pts.size() * 2 * sizeof(HF));
    for (size_t i = 0; i < pts.size(); ++i) {
        xs[i].store(&bytes[(2 * i) * sizeof(HF)]);
        ys[i].store(&bytes[(2 * i + 1) * sizeof(HF)]);
    }
    return bytes;
}

template <class F, class HostPt>
static int fixbase_run(hzk_ctx* ctx, const void* scalars, size_t n, void* out_affine, void** table_slot,
                       const HostPt& gen, const char* name) {
    HZK_CHECK(ctx, scalars && out_affine, HZK_ERR_INVALID, "%s: null pointer", name);
    if (n == 0) return HZK_OK;
    if (!*table_slot) {
        std::vector<char> t = comb_table(gen);
        void* d = nullptr;
        HZK_HIP(ctx, hipMalloc(&d, t.size()));
        HZK_HIP(ctx, hipMemcpy(d, t.data(), t.size(), hipMemcpyHostToDevice));
        *table_slot = d;
    }
    size_t need = n * sizeof(xyzz_t<F>) + n * sizeof(F) + 1024;
    int rc = ensure_scratch(ctx, need);
    if (rc) return rc;
    Arena ar{(char*)ctx->scratch, ctx->scratch_bytes};
    xyzz_t<F>* tmp = (xyzz_t<F>*)ar.take(n * sizeof(xyzz_t<F>));
    F* pre = (F*)ar.take(n * sizeof(F));
    {
        TimedLaunch tl(ctx, "fixbase");
        hipLaunchKernelGGL(k_fixbase<F>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, (const fr_t*)scalars, n,
                           (const affine_t<F>*)*table_slot, tmp);
    }
    {
        TimedLaunch tl(ctx, "batch_affine");
        size_t runs = (n + NORM_RUN - 1) / NORM_RUN;
        hipLaunchKernelGGL(k_batch_affine<F>, dim3(blocks_for(runs, 256)), dim3(256), 0, ctx->stream,
                           (const xyzz_t<F>*)tmp, n, pre, (affine_t<F>*)out_affine);
    }
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

}  // namespace hzk

using namespace hzk;

extern "C" {

int hzk_fixbase_g1(hzk_ctx* ctx, const void* scalars, size_t n, void* out_affine) {
    if (!ctx) return HZK_ERR_INVALID;
    return fixbase_run<fq_t>(ctx, scalars, n, out_affine, &ctx->comb_g1, host::g1_generator(), "fixbase_g1");
}
int hzk_fixbase_g2(hzk_ctx* ctx, const void* scalars, size_t n, void* out_affine) {
    if (!ctx) return HZK_ERR_INVALID;
    return fixbase_run<fq2_t>(ctx, scalars, n, out_affine, &ctx->comb_g2, host::g2_generator(), "fixbase_g2");
}

}  // extern "C"
