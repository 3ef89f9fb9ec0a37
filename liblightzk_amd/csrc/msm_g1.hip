// G1 instantiation of the MSM driver (kernels in msm_impl.hpp), and the G1
// MSM entry point that allocates nothing (hzk_msm_g1_noalloc).
#include <mutex>
#include "msm_impl.hpp"

namespace hzk {
int msm_g1(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest, uint32_t levels) {
    return msm_run<fq_t, host::G1x>(ctx, scalars, bases, n, win_siz, dest, levels);
}
int msm_batch_g1(hzk_ctx* ctx, const MsmJob* jobs, size_t count, int dests_on_host, int capture, int device_combine) {
    return msm_batch_run<fq_t, host::G1x>(ctx, jobs, count, dests_on_host, capture, nullptr, 0, device_combine);
}
int msm_convert_g1(hzk_ctx* ctx, const void* src, void* dst, size_t n) {
    if (!n) return 0;
    const uint64_t blocks = (n + 255) / 256, cap = (uint64_t)ctx->num_cus * 32;
    hipLaunchKernelGGL(k_msm_convert_g1_array, dim3((unsigned)(blocks < cap ? blocks : cap)), dim3(256), 0, ctx->stream,
                       (const uint4*)src, (uint4*)dst, (uint64_t)n);
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

// Scratch of hzk_msm_g1_noalloc: a file-scope device array, so a caller that
// cannot allocate (code under graph capture, serving loops, harnesses that
// forbid hipMalloc) still gets an MSM. Every GPU the library runs on holds
// its own copy until the process ends. On the host the name stands for the
// array and is no device address: kernels and copies take the address
// hipGetSymbolAddress returns for the current device, kept per device.
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)HZK_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static void* scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

// once per device, with that device current (called from hzk_ctx_create)
int msm_noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return HZK_ERR_INVALID;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_ptr[device]) return 0;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return HZK_ERR_HIP;
    }
    scratch_ptr[device] = p;
    return 0;
}

}  // namespace hzk

extern "C" int hzk_msm_g1_noalloc(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, void* dest) {
    using namespace hzk;
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, scalars && bases && dest && n > 0, HZK_ERR_INVALID, "msm_g1_noalloc: null pointer or n = 0");
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lk(scratch_mu);
        if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) p = scratch_ptr[ctx->device];
    }
    HZK_CHECK(ctx, p != nullptr, HZK_ERR_HIP, "msm_g1_noalloc: the static scratch of device %d is not mapped", ctx->device);
    MsmJob j{scalars, bases, n, dest, 0, 0};
    return msm_batch_run<fq_t, host::G1x>(ctx, &j, 1, 0, 0, p, NOALLOC_SCRATCH_BYTES);
}
