// Pippenger multi-scalar multiplication for alt_bn128 G1 and G2 on gfx950.
//
// Replaces the reference's operator (include/operator/msm.cuh:77-248: one
// counting sort of n*win_cnt (bucket, point) pairs into 2^c buckets with
// global atomics and a serial <<<1,1>>> scan, one 32-lane warp per bucket,
// fixed <<<128,1024>>> geometry valid only for c = 12) with:
//
//  1. recode   one lane per scalar: leave Montgomery form, signed c-bit digits
//              d_j in [-2^(c-1), 2^(c-1)] over W = ceil(255/c) windows (the
//              top window cannot carry because scalars are < r < 2^254), and a
//              histogram of (window, |d|) keys; digit 0 is dropped (the
//              reference accumulates bucket 0 and scales it by 0, msm.cuh:121-161).
//              Lanes of a wave that hit the same key (the 70%-{0,1} witness case
//              puts a third of all points in one bucket) are merged into one
//              atomic by a leader ballot.
//  2. scan     device-wide exclusive scan of (count, ceil(count/K)) per key ->
//              bucket offsets and level-1 segment bases; totals and the largest
//              bucket go back to the host in one small copy.
//  3. scatter  lane per scalar writes point index | sign<<31 into its bucket.
//  4. accumulate  lane per segment of <= K entries of one bucket: mixed
//              additions of affine bases into an XYZZ accumulator. Buckets of any
//              size are split into segments, so a hot bucket never serialises.
//  5. merge    strided in-place tree over each bucket's segment sums (K2-ary),
//              only as many levels as the largest bucket needs.
//  6. bucket reduction  lane per run of g buckets: running sums give
//              sum (d - lo) B_d and sum B_d; the segment's weight lo is applied
//              by double-and-add; a block per window then tree-sums segments.
//  7. window combine on the host (Horner with c doublings per window) and
//              normalisation; see host_ec.hpp for why this tail runs on the CPU.
#include <algorithm>
#include <cstring>
#include "common.hpp"
#include "curve.hpp"
#include "msm.hpp"

namespace hzk {

static constexpr uint32_t KEY_NONE = 0xffffffffu;

// ------------------------------------------------------------------ digits
template <int C>
HZK_DEV uint32_t window_bits(const fr_t& s, int j) {
    constexpr uint32_t MASK = (1u << C) - 1;
    const int lo = j * C;
    const int w = lo >> 5, sh = lo & 31;
    uint64_t v = s.l[w];
    if (w + 1 < 8) v |= (uint64_t)s.l[w + 1] << 32;
    return (uint32_t)(v >> sh) & MASK;
}

// Signed digit of window j given the incoming carry; returns |d| (0 = none).
template <int C>
HZK_DEV uint32_t signed_digit(const fr_t& s, int j, uint32_t& carry, bool& neg) {
    constexpr uint32_t NB = 1u << (C - 1);
    uint32_t raw = window_bits<C>(s, j) + carry;
    if (raw > NB) { carry = 1; neg = true; return 2 * NB - raw; }
    carry = 0; neg = false; return raw;
}

HZK_DEV uint32_t lane_id() { return __lane_id(); }

// counter[key] += 1 for every lane with key != KEY_NONE; lanes sharing the
// first active key are merged into one atomic.
HZK_DEV void wave_count(uint32_t* counter, uint32_t key) {
    uint64_t valid = __ballot(key != KEY_NONE);
    if (valid == 0) return;
    int leader = __ffsll((unsigned long long)valid) - 1;
    uint32_t lkey = __shfl(key, leader);
    uint64_t match = __ballot(key == lkey);
    if ((int)lane_id() == leader) atomicAdd(&counter[lkey], (uint32_t)__popcll(match));
    else if (key != KEY_NONE && key != lkey) atomicAdd(&counter[key], 1u);
}
// pos = atomic slot in cursor[key]; same leader merge as wave_count.
HZK_DEV uint32_t wave_slot(uint32_t* cursor, uint32_t key) {
    uint64_t valid = __ballot(key != KEY_NONE);
    if (valid == 0) return 0;
    int leader = __ffsll((unsigned long long)valid) - 1;
    uint32_t lkey = __shfl(key, leader);
    uint64_t match = __ballot(key == lkey);
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(&cursor[lkey], (uint32_t)__popcll(match));
    base = __shfl(base, leader);
    uint32_t pos = 0;
    if (key == lkey) pos = base + (uint32_t)__popcll(match & ((1ull << lane_id()) - 1));
    else if (key != KEY_NONE) pos = atomicAdd(&cursor[key], 1u);
    return pos;
}

template <int C>
__global__ void __launch_bounds__(256) k_msm_count(const fr_t* __restrict__ scalars, uint32_t n,
                                                  fr_t* __restrict__ canon, uint32_t* __restrict__ count) {
    constexpr int W = (255 + C - 1) / C;
    constexpr uint32_t NB = 1u << (C - 1);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < n;
    fr_t s = fr_t::zero();
    if (active) {
        s = scalars[i].from_mont();
        canon[i] = s;
    }
    uint32_t carry = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        bool neg;
        uint32_t d = signed_digit<C>(s, j, carry, neg);
        wave_count(count, (active && d) ? j * NB + d - 1 : KEY_NONE);
    }
}

template <int C>
__global__ void __launch_bounds__(256) k_msm_scatter(const fr_t* __restrict__ canon, uint32_t n,
                                                    uint32_t* __restrict__ cursor, uint32_t* __restrict__ sorted) {
    constexpr int W = (255 + C - 1) / C;
    constexpr uint32_t NB = 1u << (C - 1);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < n;
    fr_t s = active ? canon[i] : fr_t::zero();
    uint32_t carry = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        bool neg;
        uint32_t d = signed_digit<C>(s, j, carry, neg);
        uint32_t key = (active && d) ? j * NB + d - 1 : KEY_NONE;
        uint32_t pos = wave_slot(cursor, key);
        if (key != KEY_NONE) sorted[pos] = i | (neg ? 0x80000000u : 0u);
    }
}

// ------------------------------------------------------------------ scan
// packed value: low 32 bits = bucket count, high 32 bits = ceil(count / K)
static constexpr int SCAN_ITEMS = 4;
static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_TILE = SCAN_ITEMS * SCAN_THREADS;

HZK_DEV uint64_t pack_count(uint32_t c, uint32_t K) { return ((uint64_t)((c + K - 1) / K) << 32) | c; }

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const uint32_t* __restrict__ count, uint32_t nk, uint32_t K,
                                                             uint64_t* __restrict__ tile_sum, uint32_t* __restrict__ stats) {
    __shared__ uint64_t red[SCAN_THREADS];
    uint32_t base = blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    uint32_t mx = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        uint32_t k = base + threadIdx.x + t * SCAN_THREADS;
        if (k < nk) { uint32_t c = count[k]; s += pack_count(c, K); mx = max(mx, c); }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = SCAN_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = red[0];
    if (mx) atomicMax(&stats[2], mx);
}

// single block: exclusive scan of the tile sums in place, total to stats[0..1]
__global__ void __launch_bounds__(1024) k_scan_top(uint64_t* __restrict__ tile_sum, uint32_t ntiles, uint32_t* __restrict__ stats) {
    __shared__ uint64_t part[1024];
    uint32_t per = (ntiles + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(ntiles, b + per);
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += tile_sum[i];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partials
    for (int off = 1; off < 1024; off <<= 1) {
        uint64_t v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (uint32_t i = b; i < e; ++i) { uint64_t v = tile_sum[i]; tile_sum[i] = run; run += v; }
    if (threadIdx.x == 1023) { stats[0] = (uint32_t)part[1023]; stats[1] = (uint32_t)(part[1023] >> 32); }
}

// per tile: exclusive scan, write offsets / segment bases / cursors
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t* __restrict__ count, uint32_t nk, uint32_t K,
                                                             const uint64_t* __restrict__ tile_sum, uint32_t* __restrict__ offs,
                                                             uint32_t* __restrict__ segbase, uint32_t* __restrict__ cursor) {
    __shared__ uint64_t part[SCAN_THREADS];
    uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        uint32_t k = base + t;
        v[t] = k < nk ? pack_count(count[k], K) : 0;
        s += v[t];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint64_t x = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint64_t run = tile_sum[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        uint32_t k = base + t;
        if (k <= nk) {
            offs[k] = (uint32_t)run;
            segbase[k] = (uint32_t)(run >> 32);
            if (k < nk) cursor[k] = (uint32_t)run;
        }
        run += v[t];
    }
}

// ------------------------------------------------------------------ driver
template <int C>
static void launch_recode(hzk_ctx* ctx, const fr_t* scalars, uint32_t n, fr_t* canon, uint32_t* count) {
    TimedLaunch tl(ctx, "msm_count");
    hipLaunchKernelGGL(k_msm_count<C>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, scalars, n, canon, count);
}
template <int C>
static void launch_scatter(hzk_ctx* ctx, const fr_t* canon, uint32_t n, uint32_t* cursor, uint32_t* sorted) {
    TimedLaunch tl(ctx, "msm_scatter");
    hipLaunchKernelGGL(k_msm_scatter<C>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, canon, n, cursor, sorted);
}

typedef void (*recode_fn)(hzk_ctx*, const fr_t*, uint32_t, fr_t*, uint32_t*);
typedef void (*scatter_fn)(hzk_ctx*, const fr_t*, uint32_t, uint32_t*, uint32_t*);
template <int C>
struct CTab {
    static void fill(recode_fn* r, scatter_fn* s) {
        r[C] = launch_recode<C>;
        s[C] = launch_scatter<C>;
        CTab<C - 1>::fill(r, s);
    }
};
template <>
struct CTab<MSM_MIN_C - 1> {
    static void fill(recode_fn*, scatter_fn*) {}
};

MsmPlan msm_plan(size_t n, int force_c, int force_g) {
    MsmPlan p;
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    int c = lg - 8;
    if (c < 8) c = 8;
    if (c > 16) c = 16;
    if (force_c) c = force_c;
    p.c = c;
    p.W = (255 + c - 1) / c;
    p.NB = 1u << (c - 1);
    p.K = 64;
    uint32_t g = p.NB / 512;
    if (g < 4) g = 4;
    if (g > 16) g = 16;
    if (g > p.NB) g = p.NB;
    if (force_g) g = force_g;
    p.g = g;
    return p;
}

struct StageTables {
    recode_fn recode[MSM_MAX_C + 1];
    scatter_fn scatter[MSM_MAX_C + 1];
    StageTables() { CTab<MSM_MAX_C>::fill(recode, scatter); }
};
static const StageTables& stage_tables() {
    static StageTables t;
    return t;
}
void msm_launch_recode(hzk_ctx* ctx, int c, const void* scalars, uint32_t n, void* canon, uint32_t* count) {
    stage_tables().recode[c](ctx, (const fr_t*)scalars, n, (fr_t*)canon, count);
}
void msm_launch_scatter(hzk_ctx* ctx, int c, const void* canon, uint32_t n, uint32_t* cursor, uint32_t* sorted) {
    stage_tables().scatter[c](ctx, (const fr_t*)canon, n, cursor, sorted);
}
uint32_t msm_scan_tiles(uint32_t nk) { return (nk + SCAN_TILE) / SCAN_TILE; }   // covers index nk (the total)
void msm_launch_scan(hzk_ctx* ctx, const uint32_t* count, uint32_t nk, uint32_t K, uint64_t* tiles, uint32_t ntiles,
                     uint32_t* stats, uint32_t* offs, uint32_t* segbase, uint32_t* cursor) {
    TimedLaunch tl(ctx, "msm_scan");
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_THREADS), 0, st, count, nk, K, tiles, stats);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, tiles, ntiles, stats);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(SCAN_THREADS), 0, st, count, nk, K, tiles, offs, segbase, cursor);
}

}  // namespace hzk
