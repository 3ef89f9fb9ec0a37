// Pippenger multi-scalar multiplication for alt_bn128 G1 and G2 on gfx950.
//
// Replaces the reference's operator (include/operator/msm.cuh:77-248: one
// counting sort of n*win_cnt (bucket, point) pairs into 2^c buckets with
// global atomics and a serial <<<1,1>>> scan, one 32-lane warp per bucket,
// fixed <<<128,1024>>> geometry valid only for c = 12) with:
//
//  1. recode   one lane per scalar: leave Montgomery form, signed c-bit digits
//              d_j in [-2^(c-1), 2^(c-1)] over W = ceil(255/c) windows (the
//              top window cannot carry because scalars are < r < 2^254),
//              written window-major; digit 0 is dropped (the reference
//              accumulates bucket 0 and scales it by 0, msm.cuh:121-161).
//  2. histogram  one workgroup per (tile of scalars, window): counts in LDS
//              (2^(c-1) counters, <= 128 KiB), one coalesced row per tile. No
//              device-scope atomics: on a multi-XCD part they resolve past the
//              per-XCD L2 and ran at ~5 G/s, slower than the whole accumulation.
//  3. scan     per key a prefix over tiles, then a device-wide exclusive scan
//              of (count, ceil(count/K)) -> bucket offsets and level-1 segment
//              bases; totals and the largest bucket go back to the host in one
//              small copy.
//  4. scatter  same (tile, window) workgroups: LDS cursors seeded from the
//              scan, LDS atomics hand out slots, point index | sign<<31 is
//              written into its bucket.
//  5. accumulate  lane per segment of <= K entries of one bucket: mixed
//              additions of affine bases into an XYZZ accumulator. Buckets of any
//              size are split into segments, so a hot bucket never serialises.
//  6. merge    strided in-place tree over each bucket's segment sums (K2-ary),
//              only as many levels as the largest bucket needs.
//  7. bucket reduction  lane per run of g buckets: running sums give
//              sum (d - lo) B_d and sum B_d; the segment's weight lo is applied
//              by double-and-add; a block per window then tree-sums segments.
//  8. window combine on the host (Horner with c doublings per window) and
//              normalisation; see host_ec.hpp for why this tail runs on the CPU.
#include <algorithm>
#include <cstring>
#include "common.hpp"
#include "curve.hpp"
#include "msm.hpp"

namespace hzk {

static constexpr uint32_t DIGIT_NEG = 0x80000000u;

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

// digits[j * n + i] = |d_ij| | sign << 31, or 0 for a zero digit
template <int C>
__global__ void __launch_bounds__(256) k_msm_digits(const fr_t* __restrict__ scalars, uint32_t n,
                                                   uint32_t* __restrict__ digits) {
    constexpr int W = (255 + C - 1) / C;
    constexpr uint32_t NB = 1u << (C - 1);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t s = scalars[i].from_mont();
    uint32_t carry = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        uint32_t raw = window_bits<C>(s, j) + carry;
        uint32_t d;
        if (raw > NB) { carry = 1; d = (2 * NB - raw) | DIGIT_NEG; }
        else { carry = 0; d = raw; }
        if ((d & ~DIGIT_NEG) == 0) d = 0;
        digits[(size_t)j * n + i] = d;
    }
}

// tile t of window j covers scalars [t*tile, min(n, (t+1)*tile))
__global__ void __launch_bounds__(1024) k_msm_hist(const uint32_t* __restrict__ digits, uint32_t n, uint32_t tile,
                                                   uint32_t NB, uint32_t nk, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t cnt[];
    const uint32_t t = blockIdx.x, j = blockIdx.y;
    for (uint32_t b = threadIdx.x; b < NB; b += blockDim.x) cnt[b] = 0;
    __syncthreads();
    const uint32_t i0 = t * tile, i1 = min(n, i0 + tile);
    const uint32_t* dj = digits + (size_t)j * n;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        uint32_t d = dj[i] & ~DIGIT_NEG;
        if (d) atomicAdd(&cnt[d - 1], 1u);
    }
    __syncthreads();
    uint32_t* row = hist + (size_t)t * nk + (size_t)j * NB;
    for (uint32_t b = threadIdx.x; b < NB; b += blockDim.x) row[b] = cnt[b];
}

// per key: count = sum over tiles, hist[t][key] <- exclusive prefix over tiles
__global__ void __launch_bounds__(256) k_msm_colscan(uint32_t* __restrict__ hist, uint32_t nk, uint32_t T,
                                                     uint32_t* __restrict__ count) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    uint32_t run = 0;
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t v = hist[(size_t)t * nk + k];
        hist[(size_t)t * nk + k] = run;
        run += v;
    }
    count[k] = run;
}

__global__ void __launch_bounds__(1024) k_msm_scatter(const uint32_t* __restrict__ digits, uint32_t n, uint32_t tile,
                                                      uint32_t NB, uint32_t nk, const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ offs, uint32_t* __restrict__ sorted) {
    extern __shared__ uint32_t cur[];
    const uint32_t t = blockIdx.x, j = blockIdx.y;
    const uint32_t* row = hist + (size_t)t * nk + (size_t)j * NB;
    for (uint32_t b = threadIdx.x; b < NB; b += blockDim.x) cur[b] = offs[j * NB + b] + row[b];
    __syncthreads();
    const uint32_t i0 = t * tile, i1 = min(n, i0 + tile);
    const uint32_t* dj = digits + (size_t)j * n;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        uint32_t d = dj[i];
        uint32_t mag = d & ~DIGIT_NEG;
        if (mag) {
            uint32_t pos = atomicAdd(&cur[mag - 1], 1u);
            sorted[pos] = i | (d & DIGIT_NEG);
        }
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
    __shared__ uint32_t redmax[SCAN_THREADS];
    uint32_t base = blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    uint32_t mx = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        uint32_t k = base + threadIdx.x + t * SCAN_THREADS;
        if (k < nk) { uint32_t c = count[k]; s += pack_count(c, K); mx = max(mx, c); }
    }
    red[threadIdx.x] = s;
    redmax[threadIdx.x] = mx;
    __syncthreads();
    for (int off = SCAN_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            red[threadIdx.x] += red[threadIdx.x + off];
            redmax[threadIdx.x] = max(redmax[threadIdx.x], redmax[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tile_sum[blockIdx.x] = red[0];
        tile_sum[gridDim.x + blockIdx.x] = redmax[0];
    }
}

// single block: exclusive scan of the tile sums in place; totals and max to stats
__global__ void __launch_bounds__(1024) k_scan_top(uint64_t* __restrict__ tile_sum, uint32_t ntiles, uint32_t* __restrict__ stats) {
    __shared__ uint64_t part[1024];
    __shared__ uint32_t pmax[1024];
    uint32_t per = (ntiles + 1023) / 1024;
    uint32_t b = threadIdx.x * per, e = min(ntiles, b + per);
    uint64_t s = 0;
    uint32_t mx = 0;
    for (uint32_t i = b; i < e; ++i) { s += tile_sum[i]; mx = max(mx, (uint32_t)tile_sum[ntiles + i]); }
    part[threadIdx.x] = s;
    pmax[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint64_t v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        uint32_t m = (int)threadIdx.x >= off ? pmax[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        pmax[threadIdx.x] = max(pmax[threadIdx.x], m);
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (uint32_t i = b; i < e; ++i) { uint64_t v = tile_sum[i]; tile_sum[i] = run; run += v; }
    if (threadIdx.x == 1023) {
        stats[0] = (uint32_t)part[1023];
        stats[1] = (uint32_t)(part[1023] >> 32);
        stats[2] = pmax[1023];
    }
}

// per tile: exclusive scan, write bucket offsets and segment bases
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t* __restrict__ count, uint32_t nk, uint32_t K,
                                                             const uint64_t* __restrict__ tile_sum, uint32_t* __restrict__ offs,
                                                             uint32_t* __restrict__ segbase) {
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
        }
        run += v[t];
    }
}

// ------------------------------------------------------------------ driver pieces
template <int C>
static void launch_digits(hzk_ctx* ctx, const fr_t* scalars, uint32_t n, uint32_t* digits) {
    TimedLaunch tl(ctx, "msm_digits");
    hipLaunchKernelGGL(k_msm_digits<C>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, scalars, n, digits);
}
typedef void (*digits_fn)(hzk_ctx*, const fr_t*, uint32_t, uint32_t*);
template <int C>
struct CTab {
    static void fill(digits_fn* d) {
        d[C] = launch_digits<C>;
        CTab<C - 1>::fill(d);
    }
};
template <>
struct CTab<MSM_MIN_C - 1> {
    static void fill(digits_fn*) {}
};
struct StageTables {
    digits_fn digits[MSM_MAX_C + 1];
    StageTables() { CTab<MSM_MAX_C>::fill(digits); }
};
static const StageTables& stage_tables() {
    static StageTables t;
    return t;
}

MsmPlan msm_plan(size_t n, int force_c, int force_g, size_t lds_bytes) {
    MsmPlan p;
    int lg = 0;
    while (((size_t)1 << (lg + 1)) <= n) ++lg;
    int c = lg - 8;
    if (c < 8) c = 8;
    if (c > 16) c = 16;
    while (c > MSM_MIN_C && ((size_t)4 << (c - 1)) > lds_bytes) --c;   // histogram must fit in LDS
    if (force_c) c = force_c;
    p.c = c;
    p.W = (255 + c - 1) / c;
    p.NB = 1u << (c - 1);
    // segment length: about 2^18 lanes share the n*W additions (fewer leave
    // SIMDs idle: one lane's chain of point additions is latency-bound), and
    // at 2^24 points a typical bucket (~512 entries) stays one segment, so
    // only oversized buckets pay for the segment merge
    size_t L = n * (size_t)p.W;
    uint32_t K = 8;
    while (K < 4096 && L / (2 * K) >= ((size_t)1 << 18)) K <<= 1;
    p.K = K;
    uint32_t g = p.NB / 512;
    if (g < 4) g = 4;
    if (g > 16) g = 16;
    if (g > p.NB) g = p.NB;
    if (force_g) g = force_g;
    p.g = g;
    // scalar tiles per window for the LDS histogram / scatter workgroups
    uint32_t T = (512 + p.W - 1) / p.W;
    while (T > 1 && n / T < 4096) T >>= 1;
    p.T = T;
    p.tile = (uint32_t)((n + T - 1) / T);
    return p;
}

// Allow the histogram / scatter workgroups the LDS the largest window needs.
static void set_lds_limits() {
    static bool done = false;
    if (done) return;
    done = true;
    int bytes = (1 << (MSM_MAX_C - 1)) * 4;
    (void)hipFuncSetAttribute((const void*)k_msm_hist, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_msm_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

size_t msm_lds_bytes(const MsmPlan& p) { return (size_t)p.NB * 4; }

void msm_launch_sort(hzk_ctx* ctx, const MsmPlan& p, const void* scalars, uint32_t n, uint32_t* digits, uint32_t* hist,
                     uint32_t* count, uint64_t* tiles, uint32_t ntiles, uint32_t* stats, uint32_t* offs,
                     uint32_t* segbase) {
    const uint32_t nk = (uint32_t)p.W * p.NB;
    hipStream_t st = ctx->stream;
    set_lds_limits();
    stage_tables().digits[p.c](ctx, (const fr_t*)scalars, n, digits);
    {
        TimedLaunch tl(ctx, "msm_hist");
        hipLaunchKernelGGL(k_msm_hist, dim3(p.T, p.W), dim3(1024), p.NB * 4, st, digits, n, p.tile, p.NB, nk, hist);
    }
    {
        TimedLaunch tl(ctx, "msm_scan");
        hipLaunchKernelGGL(k_msm_colscan, dim3(blocks_for(nk, 256)), dim3(256), 0, st, hist, nk, p.T, count);
        hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_THREADS), 0, st, count, nk, p.K, tiles, stats);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, tiles, ntiles, stats);
        hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(SCAN_THREADS), 0, st, count, nk, p.K, tiles, offs, segbase);
    }
}

void msm_launch_scatter(hzk_ctx* ctx, const MsmPlan& p, const uint32_t* digits, uint32_t n, const uint32_t* hist,
                        const uint32_t* offs, uint32_t* sorted) {
    const uint32_t nk = (uint32_t)p.W * p.NB;
    TimedLaunch tl(ctx, "msm_scatter");
    hipLaunchKernelGGL(k_msm_scatter, dim3(p.T, p.W), dim3(1024), p.NB * 4, ctx->stream, digits, n, p.tile, p.NB, nk, hist,
                       offs, sorted);
}

uint32_t msm_scan_tiles(uint32_t nk) { return (nk + SCAN_TILE) / SCAN_TILE; }   // covers index nk (the total)

}  // namespace hzk
