// MSM entry points shared by the C ABI and the prover.
#pragma once
#include <cstddef>
#include <cstdint>

struct hzk_ctx;

namespace hzk {

constexpr int MSM_MIN_C = 4;
constexpr int MSM_MAX_C = 16;   // 2^(c-1) LDS counters per histogram workgroup

struct MsmPlan {
    int c;          // window bits
    int W;          // windows, ceil(255 / c)
    uint32_t NB;    // buckets per window, 2^(c-1)
    uint32_t K;     // entries per accumulation segment
    uint32_t g;     // buckets per bucket-reduction run
    uint32_t T;     // scalar tiles per window (histogram / scatter workgroups)
    uint32_t tile;  // scalars per tile
};
MsmPlan msm_plan(size_t n, int force_c, int force_g, size_t lds_bytes);

// scalar-side stages (msm.hip), shared by both groups: digits, LDS
// histograms, scans (stats[0] = entries, [1] = segments, [2] = largest bucket)
void msm_launch_sort(hzk_ctx* ctx, const MsmPlan& p, const void* scalars, uint32_t n, uint32_t* digits, uint32_t* hist,
                     uint32_t* count, uint64_t* tiles, uint32_t ntiles, uint32_t* stats, uint32_t* offs,
                     uint32_t* segbase);
void msm_launch_scatter(hzk_ctx* ctx, const MsmPlan& p, const uint32_t* digits, uint32_t n, const uint32_t* hist,
                        const uint32_t* offs, uint32_t* sorted);
uint32_t msm_scan_tiles(uint32_t nk);
size_t msm_lds_bytes(const MsmPlan& p);

int msm_g1(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest);
int msm_g2(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest);

}  // namespace hzk
