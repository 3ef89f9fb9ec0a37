// Roofline denominators measured on the device itself: the issue rate of
// v_mad_u64_u32 (the 32x32->64 integer multiply every field product is built
// from), the rate of full 256-bit Montgomery products, plain 32-bit VALU adds,
// and HBM copy bandwidth. bench.py divides achieved work by these. Kinds 4
// and 5 time a dependent chain of point doublings (G1, G2; the MSM's 29-bit
// accumulator formulas) on a single wave: the latency that a device-side
// Horner combine of the window sums (about 255 sequential doublings) pays.
#include "common.hpp"
#include "field.hpp"
#include "field29.hpp"

namespace hzk {

// 8 independent v_mad_u64_u32 chains per lane
__global__ void __launch_bounds__(256) k_probe_mad(uint64_t* out, uint32_t iters, uint32_t y) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t a[8];
    for (int j = 0; j < 8; ++j) a[j] = (uint64_t)(t * 8 + j) * 0x9e3779b97f4a7c15ULL;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint32_t x = (uint32_t)a[j];
            asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(a[j]) : "v"(x), "v"(y) : "vcc");
        }
    }
    uint64_t r = 0;
    for (int j = 0; j < 8; ++j) r ^= a[j];
    out[t] = r;
}

// 8 independent 32-bit add chains per lane
__global__ void __launch_bounds__(256) k_probe_add(uint32_t* out, uint32_t iters, uint32_t y) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t a[8];
    for (int j = 0; j < 8; ++j) a[j] = t * 8 + j;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[j]) : "v"(y));
    }
    uint32_t r = 0;
    for (int j = 0; j < 8; ++j) r ^= a[j];
    out[t] = r;
}

// 2 independent Fq Montgomery-product chains per lane
__global__ void __launch_bounds__(256) k_probe_fqmul(fq_t* out, uint32_t iters) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    fq_t x = fq_t::one(), z = fq_t::one(), y = fq_t::one();
    x.l[0] ^= t;
    z.l[1] ^= t;
    y.l[2] ^= 0x12345u;
    for (uint32_t i = 0; i < iters; ++i) {
        x = x * y;
        z = z * y;
    }
    out[t] = x + z;
}

// one wave: a chain of n doublings per lane from a lane-dependent start
// (limbs below 2^29, the top one below 2^20: a value below p)
template <class A>
__global__ void __launch_bounds__(64) k_probe_dbl(uint4* out, uint32_t n) {
    uint32_t w[A::RAW_WORDS];
    uint32_t h = 0x9e3779b9u * (threadIdx.x + 1);
    for (int i = 0; i < A::RAW_WORDS; ++i) {
        h ^= h << 13, h ^= h >> 17, h ^= h << 5;
        w[i] = (i % 9 == 8) ? (h & 0xfffffu) : (h & 0x1fffffffu);
    }
    A a = A::from_words(w);
    a.inf = false;
    for (uint32_t i = 0; i < n; ++i) a.dbl();
    a.store_raw(out + (A::RAW_WORDS / 4) * threadIdx.x);
}

__global__ void __launch_bounds__(256) k_probe_copy(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

}  // namespace hzk

using namespace hzk;

extern "C" int hzk_roofline_probe(hzk_ctx* ctx, int kind, double* value) {
    if (!ctx || !value) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, kind >= 0 && kind <= 5, HZK_ERR_INVALID, "probe: kind %d", kind);
    hipDeviceProp_t prop;
    HZK_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    const uint32_t blocks = prop.multiProcessorCount * 32;   // 32 waves of 64 lanes... per CU: 8 per SIMD
    const uint32_t threads = 256;
    const size_t lanes = (size_t)blocks * threads;
    size_t copy_bytes = (size_t)1 << 30;
    size_t need = kind == 3 ? 2 * copy_bytes : lanes * sizeof(fq_t);
    int rc = ensure_scratch(ctx, need);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HZK_HIP(ctx, hipEventCreate(&e0));
    HZK_HIP(ctx, hipEventCreate(&e1));
    double work = 0;
    for (int rep = 0; rep < 2; ++rep) {   // first pass warms up
        HZK_HIP(ctx, hipEventRecord(e0, ctx->stream));
        if (kind == 0) {
            const uint32_t iters = 2048;
            hipLaunchKernelGGL(k_probe_mad, dim3(blocks), dim3(threads), 0, ctx->stream, (uint64_t*)ctx->scratch, iters, 0x9e3779b9u);
            work = (double)lanes * iters * 8;
        } else if (kind == 1) {
            const uint32_t iters = 128;
            hipLaunchKernelGGL(k_probe_fqmul, dim3(blocks), dim3(threads), 0, ctx->stream, (fq_t*)ctx->scratch, iters);
            work = (double)lanes * iters * 2;
        } else if (kind == 2) {
            const uint32_t iters = 8192;
            hipLaunchKernelGGL(k_probe_add, dim3(blocks), dim3(threads), 0, ctx->stream, (uint32_t*)ctx->scratch, iters, 3u);
            work = (double)lanes * iters * 8;
        } else if (kind >= 4) {
            const uint32_t n = 256;
            if (kind == 4) hipLaunchKernelGGL(k_probe_dbl<g1acc29>, dim3(1), dim3(64), 0, ctx->stream, (uint4*)ctx->scratch, n);
            else hipLaunchKernelGGL(k_probe_dbl<g2acc29>, dim3(1), dim3(64), 0, ctx->stream, (uint4*)ctx->scratch, n);
            work = n;   // doublings along one lane's chain
        } else {
            size_t n = copy_bytes / 16;
            hipLaunchKernelGGL(k_probe_copy, dim3(prop.multiProcessorCount * 16), dim3(threads), 0, ctx->stream,
                               (const uint4*)ctx->scratch, (uint4*)((char*)ctx->scratch + copy_bytes), n);
            work = 2.0 * copy_bytes;
        }
        HZK_HIP(ctx, hipGetLastError());
        HZK_HIP(ctx, hipEventRecord(e1, ctx->stream));
        HZK_HIP(ctx, hipEventSynchronize(e1));
    }
    float ms = 0;
    HZK_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *value = work / (ms * 1e-3);   // ops/s or bytes/s
    return HZK_OK;
}
