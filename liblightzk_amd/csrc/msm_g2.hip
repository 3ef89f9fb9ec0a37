// G2 instantiation of the MSM driver (kernels in msm_impl.hpp).
#include "msm_impl.hpp"

namespace hzk {
int msm_g2(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest, uint32_t levels) {
    return msm_run<fq2_t, host::G2x>(ctx, scalars, bases, n, win_siz, dest, levels);
}
int msm_batch_g2(hzk_ctx* ctx, const MsmJob* jobs, size_t count, int dests_on_host, int capture, int device_combine) {
    return msm_batch_run<fq2_t, host::G2x>(ctx, jobs, count, dests_on_host, capture, nullptr, 0, device_combine);
}
int msm_convert_g2(hzk_ctx* ctx, const void* src, void* dst, size_t n) {
    if (!n) return 0;
    const uint64_t blocks = (n + 255) / 256, cap = (uint64_t)ctx->num_cus * 32;
    hipLaunchKernelGGL(k_msm_convert_g2_array, dim3((unsigned)(blocks < cap ? blocks : cap)), dim3(256), 0, ctx->stream,
                       (const uint4*)src, (uint4*)dst, (uint64_t)n);
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}
}  // namespace hzk
