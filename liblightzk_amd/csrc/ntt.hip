// Fr number-theoretic transforms on gfx950: NTT, inverse NTT and coset NTT.
//
// Replaces number_theory_transformation (include/operator/poly.cuh:65-73:
// one launch and one full HBM read+write per radix-2 stage, half the lanes
// idle, then a bit-reverse launch and a scaling launch) while keeping its
// semantics exactly: natural-order input and output,
//   out[k] = factor * sum_j in[j] * w^(j k),   w = omega_powers[1],
// with omega_powers[e] = w^e for e < deg/2 (the table layout of ntt_setup,
// include/protocol/groth16-operator.cuh:6-32) and the same decimation-in-
// frequency butterfly (a, b) -> (a + b, (a - b) w^((j mod h) 2^s)).
//
// Here the log2(deg) stages are grouped into passes: each workgroup stages a
// 2^11-element tile (64 KiB) in LDS — 2^R butterfly rows by G = 2^(11-R)
// columns that are contiguous in HBM, so loads and stores stay coalesced —
// runs R stages there, and writes the tile back. 2^20 takes two passes
// (R = 9, then the last 11 stages, whose rows are contiguous). One more pass
// applies the bit-reverse permutation fused with the scale factor.
#include <vector>
#include "common.hpp"
#include "field.hpp"
#include "host_ec.hpp"

namespace hzk {

static constexpr int NTT_TILE_BITS = 11;
static constexpr int NTT_THREADS = 256;
static constexpr int NTT_MAX_LOG = 26;

// one pass of stages [s0, s0 + R) over tiles of 2^R rows x 2^gbits columns
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(fr_t* __restrict__ x, const fr_t* __restrict__ tw, int k, int s0,
                                                          int R, int gbits) {
    extern __shared__ fr_t sh[];
    const int low = k - s0 - R;
    const uint32_t G = 1u << gbits;
    const uint32_t T = 1u << (R + gbits);
    const uint32_t lchunks = (1u << low) >> gbits;
    const uint32_t b = blockIdx.x;
    const uint32_t hi = b / lchunks, lc = b % lchunks;
    const size_t base = ((size_t)hi << (k - s0)) | ((size_t)lc << gbits);
    for (uint32_t e = threadIdx.x; e < T; e += NTT_THREADS) {
        uint32_t m = e >> gbits, g = e & (G - 1);
        sh[e] = x[base + ((size_t)m << low) + g];
    }
    __syncthreads();
    for (int st = 0; st < R; ++st) {
        const int s = s0 + st;
        const int mp = R - 1 - st;
        const uint32_t mmask = (1u << mp) - 1;
        for (uint32_t bf = threadIdx.x; bf < T / 2; bf += NTT_THREADS) {
            uint32_t g = bf & (G - 1), q = bf >> gbits;
            uint32_t m0 = ((q >> mp) << (mp + 1)) | (q & mmask);
            uint32_t i0 = (m0 << gbits) | g, i1 = ((m0 | (1u << mp)) << gbits) | g;
            size_t e = ((((size_t)(m0 & mmask)) << low) | ((size_t)lc << gbits | g)) << s;
            fr_t a = sh[i0], c = sh[i1];
            sh[i0] = a + c;
            sh[i1] = (a - c) * tw[e];
        }
        __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < T; e += NTT_THREADS) {
        uint32_t m = e >> gbits, g = e & (G - 1);
        x[base + ((size_t)m << low) + g] = sh[e];
    }
}

// bit-reverse permutation (k bits) fused with an optional scale factor
__global__ void __launch_bounds__(256) k_bitrev_scale(fr_t* __restrict__ x, size_t n, int k, const fr_t* __restrict__ factor) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t r = (size_t)(__brev((uint32_t)i) >> (32 - k));
    if (i > r) return;
    fr_t a = x[i], c = x[r];
    if (factor) {
        fr_t f = *factor;
        a = a * f;
        c = c * f;
    }
    x[i] = c;
    x[r] = a;
}

// x[j] *= tab[j] (* scale)
__global__ void __launch_bounds__(256) k_mul_table(fr_t* __restrict__ x, const fr_t* __restrict__ tab, size_t n,
                                                   const fr_t* __restrict__ scale) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t v = x[i] * tab[i];
    if (scale) v = v * *scale;
    x[i] = v;
}

// out[e] = w^e for e < n, from sq[b] = w^(2^b)
__global__ void __launch_bounds__(256) k_powers(fr_t* __restrict__ out, size_t n, const fr_t* __restrict__ sq) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    fr_t acc = fr_t::one();
    for (int b = 0; (e >> b) != 0; ++b)
        if ((e >> b) & 1) acc = acc * sq[b];
    out[e] = acc;
}

// ------------------------------------------------------------------ host side
// Fr constants in Montgomery form: primitive 2^k-th roots of unity
// w_k = 5^((r-1)/2^k) (5 generates Fr*, alt_bn128_init.cu:50) and inverses
// are derived on the host from w_28, the value the reference stores as
// root_of_unity (alt_bn128_init.cu:51).
static host::Fr fr_from_u64(uint64_t v) {
    host::Fr t{{v, 0, 0, 0}};
    host::Fr r2;
    std::memcpy(r2.v, host::FR.r2, 32);
    return t * r2;
}
static host::Fr root_of_unity(int k) {
    // w_28 = 5^t, t = (r-1)/2^28
    static const uint64_t t[4] = {0x9b9709143e1f593fULL, 0x181585d2833e8487ULL, 0x131a029b85045b68ULL, 0x000000030644e72eULL};
    host::Fr w = fr_from_u64(5).pow(t);
    for (int i = 28; i > k; --i) w = w.sqr();
    return w;
}

// per transform size: forward / inverse twiddles (n/2 each), coset powers
// g^j and g^-j (n each) and 1/n; built on first use, kept for the context
struct NttTables {
    fr_t* tw[2] = {nullptr, nullptr};
    fr_t* coset[2] = {nullptr, nullptr};
    fr_t* ninv = nullptr;
};
struct NttCache {
    NttTables t[NTT_MAX_LOG + 1];
};

static NttCache& cache_of(hzk_ctx* ctx) {
    if (!ctx->ntt_cache) ctx->ntt_cache = new NttCache();
    return *(NttCache*)ctx->ntt_cache;
}

void ntt_release(hzk_ctx* ctx) {
    if (!ctx->ntt_cache) return;
    NttCache& c = *(NttCache*)ctx->ntt_cache;
    for (auto& t : c.t)
        for (auto p : {t.tw[0], t.tw[1], t.coset[0], t.coset[1], t.ninv})
            if (p) (void)hipFree(p);
    delete &c;
    ctx->ntt_cache = nullptr;
}

static int fill_powers(hzk_ctx* ctx, fr_t* out, size_t n, host::Fr w) {
    std::vector<host::Fr> sq(64);
    for (int b = 0; b < 64; ++b) { sq[b] = w; w = w.sqr(); }
    fr_t* dsq = nullptr;
    HZK_HIP(ctx, hipMalloc(&dsq, 64 * 32));
    HZK_HIP(ctx, hipMemcpy(dsq, sq.data(), 64 * 32, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_powers, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, out, n, (const fr_t*)dsq);
    HZK_HIP(ctx, hipGetLastError());
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HZK_HIP(ctx, hipFree(dsq));
    return 0;
}

static int ensure_tables(hzk_ctx* ctx, int logn, NttTables** out) {
    NttTables& t = cache_of(ctx).t[logn];
    *out = &t;
    if (t.ninv) return 0;
    size_t n = (size_t)1 << logn;
    size_t half = n / 2 ? n / 2 : 1;
    host::Fr w = root_of_unity(logn), wi = w.inverse();
    host::Fr g = fr_from_u64(5), gi = g.inverse();
    host::Fr ninv = fr_from_u64(n).inverse();
    for (int d = 0; d < 2; ++d) {
        HZK_HIP(ctx, hipMalloc(&t.tw[d], half * 32));
        HZK_HIP(ctx, hipMalloc(&t.coset[d], n * 32));
        int rc = fill_powers(ctx, t.tw[d], half, d ? wi : w);
        if (rc) return rc;
        rc = fill_powers(ctx, t.coset[d], n, d ? gi : g);
        if (rc) return rc;
    }
    fr_t* ni = nullptr;
    HZK_HIP(ctx, hipMalloc(&ni, 32));
    HZK_HIP(ctx, hipMemcpy(ni, ninv.v, 32, hipMemcpyHostToDevice));
    t.ninv = ni;
    return 0;
}

static void set_ntt_lds() {
    static bool done = false;
    if (done) return;
    done = true;
    (void)hipFuncSetAttribute((const void*)k_ntt_pass, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (1 << NTT_TILE_BITS) * 32);
}

// In-place transform with twiddle table tw (w^e, e < n/2), optional device factor.
int ntt_run(hzk_ctx* ctx, fr_t* x, int k, const fr_t* tw, const fr_t* factor) {
    size_t n = (size_t)1 << k;
    hipStream_t st = ctx->stream;
    set_ntt_lds();
    if (k == 0) {
        if (factor) hipLaunchKernelGGL(k_bitrev_scale, dim3(1), dim3(1), 0, st, x, n, 1, factor);   // 1 x 1: scale only
        return 0;
    }
    // pass plan: the last pass takes the lowest min(k, 11) stages (contiguous
    // rows); the rest split evenly into passes of <= 9 stages (>= 4 columns)
    int last = k < NTT_TILE_BITS ? k : NTT_TILE_BITS;
    int rest = k - last;
    int npass = (rest + 8) / 9;
    int s0 = 0;
    {
        TimedLaunch tl(ctx, "ntt_pass");
        for (int p = 0; p < npass; ++p) {
            int R = rest / npass + (p < rest % npass ? 1 : 0);
            int gbits = NTT_TILE_BITS - R;
            int low = k - s0 - R;
            if (gbits > low) gbits = low;
            size_t tiles = n >> (R + gbits);
            hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)tiles), dim3(NTT_THREADS), (32u << (R + gbits)), st, x, tw, k, s0, R,
                               gbits);
            s0 += R;
        }
        hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)(n >> last)), dim3(NTT_THREADS), (32u << last), st, x, tw, k, s0, last, 0);
    }
    {
        TimedLaunch tl(ctx, "ntt_bitrev");
        hipLaunchKernelGGL(k_bitrev_scale, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, k, factor);
    }
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

static int log2_exact(size_t n) {
    int k = 0;
    while (((size_t)1 << k) < n) ++k;
    return ((size_t)1 << k) == n ? k : -1;
}

int ntt_standard(hzk_ctx* ctx, fr_t* x, size_t deg, int inverse) {
    int k = log2_exact(deg);
    HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG && k <= 28, HZK_ERR_INVALID, "ntt: deg %zu is not a power of two <= 2^%d", deg,
              NTT_MAX_LOG);
    NttTables* t;
    int rc = ensure_tables(ctx, k, &t);
    if (rc) return rc;
    return ntt_run(ctx, x, k, t->tw[inverse ? 1 : 0], inverse ? t->ninv : nullptr);
}

int coset_ntt(hzk_ctx* ctx, fr_t* x, size_t deg, int inverse) {
    int k = log2_exact(deg);
    HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "coset_ntt: deg %zu is not a power of two <= 2^%d", deg,
              NTT_MAX_LOG);
    NttTables* t;
    int rc = ensure_tables(ctx, k, &t);
    if (rc) return rc;
    if (!inverse) {
        {
            TimedLaunch tl(ctx, "ntt_coset_scale");
            hipLaunchKernelGGL(k_mul_table, dim3(blocks_for(deg, 256)), dim3(256), 0, ctx->stream, x, (const fr_t*)t->coset[0],
                               deg, (const fr_t*)nullptr);
        }
        return ntt_run(ctx, x, k, t->tw[0], nullptr);
    }
    rc = ntt_run(ctx, x, k, t->tw[1], t->ninv);
    if (rc) return rc;
    TimedLaunch tl(ctx, "ntt_coset_scale");
    hipLaunchKernelGGL(k_mul_table, dim3(blocks_for(deg, 256)), dim3(256), 0, ctx->stream, x, (const fr_t*)t->coset[1], deg,
                       (const fr_t*)nullptr);
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

int ntt_standard_raw(hzk_ctx* ctx, void* x, size_t deg, int inverse) { return ntt_standard(ctx, (fr_t*)x, deg, inverse); }
int coset_ntt_raw(hzk_ctx* ctx, void* x, size_t deg, int inverse) { return coset_ntt(ctx, (fr_t*)x, deg, inverse); }

}  // namespace hzk

using namespace hzk;

extern "C" {

int hzk_ntt(hzk_ctx* ctx, void* coeffs, size_t deg, const void* omega_powers, int inverse, const void* factor) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, coeffs != nullptr, HZK_ERR_INVALID, "ntt: null coefficients");
    if (!omega_powers) {
        if (factor) {
            int k = log2_exact(deg);
            HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "ntt: deg %zu invalid", deg);
            NttTables* t;
            int rc = ensure_tables(ctx, k, &t);
            if (rc) return rc;
            return ntt_run(ctx, (fr_t*)coeffs, k, t->tw[inverse ? 1 : 0], (const fr_t*)factor);
        }
        return ntt_standard(ctx, (fr_t*)coeffs, deg, inverse);
    }
    int k = log2_exact(deg);
    HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "ntt: deg %zu is not a power of two <= 2^%d", deg, NTT_MAX_LOG);
    return ntt_run(ctx, (fr_t*)coeffs, k, (const fr_t*)omega_powers, (const fr_t*)factor);
}

int hzk_coset_ntt(hzk_ctx* ctx, void* coeffs, size_t deg, int inverse) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, coeffs != nullptr, HZK_ERR_INVALID, "coset_ntt: null coefficients");
    return coset_ntt(ctx, (fr_t*)coeffs, deg, inverse);
}

}  // extern "C"
