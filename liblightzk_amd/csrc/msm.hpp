// MSM entry points shared by the C ABI and the prover.
#pragma once
#include <cstddef>
#include <cstdint>

struct hzk_ctx;

namespace hzk {

constexpr int MSM_MIN_C = 4;
constexpr int MSM_MAX_C = 20;

struct MsmPlan {
    int c;          // window bits
    int W;          // windows, ceil(255 / c)
    uint32_t NB;    // buckets per window, 2^(c-1)
    uint32_t K;     // entries per accumulation segment
    uint32_t g;     // buckets per bucket-reduction run
};
MsmPlan msm_plan(size_t n, int force_c, int force_g);

// scalar-side stages (msm.hip), shared by both groups
void msm_launch_recode(hzk_ctx* ctx, int c, const void* scalars, uint32_t n, void* canon, uint32_t* count);
void msm_launch_scatter(hzk_ctx* ctx, int c, const void* canon, uint32_t n, uint32_t* cursor, uint32_t* sorted);
uint32_t msm_scan_tiles(uint32_t nk);
void msm_launch_scan(hzk_ctx* ctx, const uint32_t* count, uint32_t nk, uint32_t K, uint64_t* tiles, uint32_t ntiles,
                     uint32_t* stats, uint32_t* offs, uint32_t* segbase, uint32_t* cursor);

int msm_g1(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest);
int msm_g2(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest);

}  // namespace hzk
