// Pippenger multi-scalar multiplication for alt_bn128 G1 and G2 on gfx950:
// scalar-side stages, batched over any number of MSM instances.
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
//  3. scan     per key a prefix over tiles, then an exclusive scan of the
//              bucket counts -> bucket offsets; per-instance totals (entries,
//              chunks, largest bucket) land in device stats, so no stage waits
//              for the host.
//  4. scatter  same (tile, window) workgroups: LDS cursors seeded from the
//              scan, LDS atomics hand out slots, point index | sign<<31 is
//              written into its bucket.
//  5-8 (msm_impl.hpp) chunked accumulation with boundary fix-up, bucket
//              reduction, window sums;
//  9. window combine on the host (Horner with c doublings per window) and
//              normalisation; see host_ec.hpp for why this tail runs on the CPU.
//
// Every stage is one launch over the whole batch: each instance owns a
// contiguous range of every work space (MsmSpace), and a lane finds its
// instance by binary search over the range starts. A single MSM is a batch of
// one; a batch of uneven sizes fills the GPU where one small MSM cannot.
#include <algorithm>
#include <cstring>
#include <mutex>
#include "common.hpp"
#include "curve.hpp"
#include "msm.hpp"

namespace hzk {

static constexpr uint32_t DIGIT_NEG = 0x80000000u;
static constexpr uint32_t MSM_MAX_FINE = (1u << (MSM_MAX_C - 1)) / MSM_COARSE;   // buckets per coarse group at most
static constexpr int SCAN_ITEMS = 4;
static constexpr int SCAN_THREADS = 256;
static_assert(SCAN_ITEMS * SCAN_THREADS == MSM_SCAN_TILE, "scan tile");

// largest i < ninst with beg[i] <= gid (beg[0] = 0 <= gid < beg[ninst])
HZK_DEV uint32_t find_inst(const uint64_t* __restrict__ beg, uint32_t ninst, uint64_t gid) {
    uint32_t lo = 0, hi = ninst;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (beg[mid] <= gid) lo = mid; else hi = mid;
    }
    return lo;
}
HZK_DEV const uint64_t* space(const MsmBatchDev& b, int s) { return b.begin + (size_t)s * (b.ninst + 1); }

// ------------------------------------------------------------------ 1. digits
// digits[j * n + i] = |d_ij| | sign << 31, or 0 for a zero digit. For a level
// table of Lg levels the same array is, window by window of the pipeline, the
// digits of Lg consecutive digit windows over nd * Lg virtual points
// 2^(c (j mod Lg)) P_i, so Lg digit windows share one bucket set.
__global__ void __launch_bounds__(256) k_msm_digits(MsmBatchDev b, uint32_t* __restrict__ digits) {
    __shared__ uint32_t limbs[256][9];
    const uint64_t* beg = space(b, SP_DIGIT);
    const uint64_t* ebeg = space(b, SP_ENTRY);
    const uint64_t total = beg[b.ninst];
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t ii = find_inst(beg, b.ninst, gid);
        const MsmInst& d = b.inst[ii];
        uint32_t idx = (uint32_t)(gid - beg[ii]);
        fr_t s = ((const fr_t*)d.scalars)[idx].from_mont();
        uint32_t* my = limbs[threadIdx.x];
        for (int w = 0; w < 8; ++w) my[w] = s.l[w];
        my[8] = 0;
        uint32_t carry = 0;
        uint32_t* out = digits + ebeg[ii] + idx;
        // window j: bits [lo, lo + wj), wj = wb + 1 for j < wx, then wb (<= c);
        // digits in [-2^(wj-1), 2^(wj-1)]
        uint32_t lo = 0, wj = d.wb + (d.wx ? 1u : 0u);
        uint32_t half = 1u << (wj - 1), mask = (1u << wj) - 1;
        for (uint32_t j = 0; j < d.W * d.Lg; ++j) {   // level tables: windows past Wd pad the last set
            uint32_t dg = 0;
            if (j == d.wx && d.wx) {
                wj = d.wb;
                half = 1u << (wj - 1);
                mask = (1u << wj) - 1;
            }
            if (j < d.Wd) {
                uint32_t w = lo >> 5, sh = lo & 31;
                uint64_t v = (uint64_t)my[w] | ((uint64_t)my[w + 1] << 32);
                uint32_t raw = ((uint32_t)(v >> sh) & mask) + carry;
                if (raw > half) { carry = 1; dg = (2 * half - raw) | DIGIT_NEG; }
                else { carry = 0; dg = raw; }
                if ((dg & ~DIGIT_NEG) == 0) dg = 0;
                lo += wj;
            }
            out[(size_t)j * d.nd] = dg;   // window j of scalar idx (level tables: set j / Lg, point (j % Lg) nd + idx)
        }
    }
}

// ------------------------------------------------------------------ 2. histogram
// workgroup (t, j) of an instance counts the digits of scalars [t*tile, (t+1)*tile) in window j
__global__ void __launch_bounds__(1024) k_msm_hist(MsmBatchDev b, const uint32_t* __restrict__ digits, uint32_t* __restrict__ hist) {
    extern __shared__ uint32_t cnt[];
    const uint64_t* beg = space(b, SP_TILEWG);
    uint32_t ii = find_inst(beg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    uint32_t local = (uint32_t)(blockIdx.x - beg[ii]);
    uint32_t t = local % d.T, j = local / d.T;
    for (uint32_t x = threadIdx.x; x < d.NB; x += blockDim.x) cnt[x] = 0;
    __syncthreads();
    const uint32_t i0 = t * d.tile, i1 = min(d.n, i0 + d.tile);
    const size_t abs0 = space(b, SP_ENTRY)[ii] + (size_t)j * d.n;
    const uint32_t* dj = digits + abs0;
    auto count = [&](uint32_t dg) {
        const uint32_t m = dg & ~DIGIT_NEG;
        if (m) atomicAdd(&cnt[m - 1], 1u);
    };
    // 16-byte loads (four digits per lane) over the aligned middle of the
    // tile, single digits for its at most 3 + 3 unaligned ends
    uint32_t ia = i0 + ((4u - (uint32_t)((abs0 + i0) & 3)) & 3u);
    if (ia > i1) ia = i1;
    const uint32_t nv = (i1 - ia) / 4, ib = ia + 4 * nv;
    for (uint32_t i = i0 + threadIdx.x; i < ia; i += blockDim.x) count(dj[i]);
    for (uint32_t i = ib + threadIdx.x; i < i1; i += blockDim.x) count(dj[i]);
    const uint4* dv = (const uint4*)(dj + ia);
    for (uint32_t q = threadIdx.x; q < nv; q += blockDim.x) {
        const uint4 w = dv[q];
        count(w.x);
        count(w.y);
        count(w.z);
        count(w.w);
    }
    __syncthreads();
    uint32_t* row = hist + space(b, SP_HIST)[ii] + (size_t)t * d.nk + (size_t)j * d.NB;
    for (uint32_t x = threadIdx.x; x < d.NB; x += blockDim.x) row[x] = cnt[x];
}

// ------------------------------------------------------------------ 3. scans
// per key: count = sum over tiles; hist[t][key] <- exclusive prefix over tiles
__global__ void __launch_bounds__(256) k_msm_colscan(MsmBatchDev b, uint32_t* __restrict__ hist, uint32_t* __restrict__ count) {
    const uint64_t* beg = space(b, SP_KEY);
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= beg[b.ninst]) return;
    uint32_t ii = find_inst(beg, b.ninst, gid);
    const MsmInst& d = b.inst[ii];
    uint32_t k = (uint32_t)(gid - beg[ii]);
    if (k >= d.nk) { count[gid] = 0; return; }
    uint32_t* h = hist + space(b, SP_HIST)[ii] + k;
    uint32_t run = 0;
    for (uint32_t t = 0; t < d.T; ++t) {
        uint32_t v = h[(size_t)t * d.nk];
        h[(size_t)t * d.nk] = run;
        run += v;
    }
    count[gid] = run;
}

// one block per scan tile: tile sum and tile maximum
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(MsmBatchDev b, const uint32_t* __restrict__ count,
                                                             uint64_t* __restrict__ tile_sum, uint64_t* __restrict__ tile_max) {
    __shared__ uint64_t red[SCAN_THREADS];
    __shared__ uint32_t redmax[SCAN_THREADS];
    const uint64_t* tbeg = space(b, SP_SCANTILE);
    uint32_t ii = find_inst(tbeg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    uint64_t kbase = space(b, SP_KEY)[ii];
    uint32_t base = (uint32_t)(blockIdx.x - tbeg[ii]) * MSM_SCAN_TILE;
    uint64_t s = 0;
    uint32_t mx = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        uint32_t k = base + threadIdx.x + t * SCAN_THREADS;
        if (k <= d.nk) { uint32_t c = count[kbase + k]; s += c; mx = max(mx, c); }
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
        tile_max[blockIdx.x] = redmax[0];
    }
}

// one block per instance: exclusive scan of its tile sums; totals to stats
__global__ void __launch_bounds__(1024) k_scan_top(MsmBatchDev b, uint64_t* __restrict__ tile_sum, const uint64_t* __restrict__ tile_max) {
    __shared__ uint64_t part[1024];
    __shared__ uint32_t pmax[1024];
    const uint64_t* tbeg = space(b, SP_SCANTILE);
    const uint32_t ii = blockIdx.x;
    const uint64_t t0 = tbeg[ii];
    const uint32_t ntiles = (uint32_t)(tbeg[ii + 1] - t0);
    uint32_t per = (ntiles + 1023) / 1024;
    uint32_t lo = threadIdx.x * per, hi = min(ntiles, lo + per);
    uint64_t s = 0;
    uint32_t mx = 0;
    for (uint32_t i = lo; i < hi; ++i) { s += tile_sum[t0 + i]; mx = max(mx, (uint32_t)tile_max[t0 + i]); }
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
    for (uint32_t i = lo; i < hi; ++i) { uint64_t v = tile_sum[t0 + i]; tile_sum[t0 + i] = run; run += v; }
    if (threadIdx.x == 1023) {
        const uint32_t entries = (uint32_t)part[1023], K = b.inst[ii].K;
        b.stats[4 * ii + 0] = entries;
        b.stats[4 * ii + 1] = (entries + K - 1) / K;
        b.stats[4 * ii + 2] = pmax[1023];
        b.stats[4 * ii + 3] = 0;
    }
}

// per tile: exclusive scan -> bucket offsets for keys 0..nk
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(MsmBatchDev b, const uint32_t* __restrict__ count,
                                                             const uint64_t* __restrict__ tile_sum, uint32_t* __restrict__ offs) {
    __shared__ uint64_t part[SCAN_THREADS];
    const uint64_t* tbeg = space(b, SP_SCANTILE);
    uint32_t ii = find_inst(tbeg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    uint64_t kbase = space(b, SP_KEY)[ii];
    uint32_t base = (uint32_t)(blockIdx.x - tbeg[ii]) * MSM_SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        uint32_t k = base + t;
        v[t] = k <= d.nk ? count[kbase + k] : 0;
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
        if (k <= d.nk) offs[kbase + k] = (uint32_t)run;
        run += v[t];
    }
}

// ------------------------------------------------------------------ 4. scatter
__global__ void __launch_bounds__(1024) k_msm_scatter(MsmBatchDev b, const uint32_t* __restrict__ digits,
                                                      const uint32_t* __restrict__ hist, const uint32_t* __restrict__ offs,
                                                      uint32_t* __restrict__ sorted) {
    extern __shared__ uint32_t cur[];
    const uint64_t* beg = space(b, SP_TILEWG);
    uint32_t ii = find_inst(beg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    if (d.tp) return;   // scattered in two passes below
    uint32_t local = (uint32_t)(blockIdx.x - beg[ii]);
    uint32_t t = local % d.T, j = local / d.T;
    const uint32_t* row = hist + space(b, SP_HIST)[ii] + (size_t)t * d.nk + (size_t)j * d.NB;
    const uint32_t* of = offs + space(b, SP_KEY)[ii] + (size_t)j * d.NB;
    for (uint32_t x = threadIdx.x; x < d.NB; x += blockDim.x) cur[x] = of[x] + row[x];
    __syncthreads();
    const uint32_t i0 = t * d.tile, i1 = min(d.n, i0 + d.tile);
    const uint64_t ebase = space(b, SP_ENTRY)[ii];
    const uint32_t* dj = digits + ebase + (size_t)j * d.n;
    uint32_t* out = sorted + ebase;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        uint32_t dg = dj[i];
        uint32_t m = dg & ~DIGIT_NEG;
        if (m) {
            uint32_t pos = atomicAdd(&cur[m - 1], 1u);
            out[pos] = i | (dg & DIGIT_NEG);
        }
    }
}

// ------------------------------------------------------------------ 4'. two-pass scatter
// With 2^(c-1) >= 2048 buckets a single scatter writes each 4-byte entry to a
// different bucket range: on the MI355X that became ~8x the entry bytes in HBM
// writes (every write a separate partially-filled sector). The two passes
// below write coalesced runs instead: pass 1 partitions each tile's entries by
// coarse key (the top 8 bits of the bucket index, 256 groups per window) into
// an 8-byte staging array laid out exactly like the final bucket ranges;
// pass 2 sorts each coarse group by its low bits. In both, a round of
// SC_ROUND entries is counting-sorted in LDS first, so consecutive lanes write
// consecutive addresses of one group. 8192-entry rounds give a coarse group
// about 32 entries (128 B) per round at 2^24 points; 4096 left half-filled
// lines (2^24 MSM: both passes 2.18 ms with 4096, 1.84 ms with 8192, slower
// again with 16384 and its single workgroup per CU).
static constexpr uint32_t SC_ROUND = 8192;
static constexpr uint32_t SC_THREADS = 1024;
static constexpr uint32_t SC_PER = SC_ROUND / SC_THREADS;
static constexpr uint32_t SC_NONE = 0xffffffffu;

// exclusive scan of n <= 1024 LDS counters (one thread per counter)
HZK_DEV void lds_exclusive_scan(const uint32_t* cnt, uint32_t* out, uint32_t* tmp, uint32_t n) {
    const uint32_t t = threadIdx.x;
    if (t < n) tmp[t] = cnt[t];
    __syncthreads();
    for (uint32_t off = 1; off < n; off <<= 1) {
        uint32_t v = (t < n && t >= off) ? tmp[t - off] : 0;
        __syncthreads();
        if (t < n) tmp[t] += v;
        __syncthreads();
    }
    if (t < n) out[t] = tmp[t] - cnt[t];
    __syncthreads();
}

// pass 1: workgroup per (tile, window) of a two-pass instance
__global__ void __launch_bounds__(SC_THREADS) k_msm_scatter_coarse(MsmBatchDev b, const uint32_t* __restrict__ digits,
                                                                   const uint32_t* __restrict__ hist, const uint32_t* __restrict__ offs,
                                                                   uint2* __restrict__ coarse) {
    __shared__ uint32_t cur[MSM_COARSE], cnt[MSM_COARSE], lstart[MSM_COARSE], tmp[MSM_COARSE];
    __shared__ uint2 stage[SC_ROUND];
    __shared__ uint8_t skey[SC_ROUND];
    const uint64_t* beg = space(b, SP_TILEWG);
    const uint32_t ii = find_inst(beg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    if (!d.tp) return;
    const uint32_t local = (uint32_t)(blockIdx.x - beg[ii]);
    const uint32_t t = local % d.T, j = local / d.T;
    const uint32_t FB = d.NB / MSM_COARSE, fshift = __ffs(FB) - 1;
    const uint32_t tid = threadIdx.x;
    // cursor of coarse group h: start of its first bucket + this tile's prefix over its buckets
    if (tid < MSM_COARSE) {
        const uint32_t* row = hist + space(b, SP_HIST)[ii] + (size_t)t * d.nk + (size_t)j * d.NB + (size_t)tid * FB;
        uint32_t acc = offs[space(b, SP_KEY)[ii] + (size_t)j * d.NB + (size_t)tid * FB];
        for (uint32_t q = 0; q < FB; ++q) acc += row[q];
        cur[tid] = acc;
    }
    const uint32_t i0 = t * d.tile, i1 = min(d.n, i0 + d.tile);
    const uint64_t ebase = space(b, SP_ENTRY)[ii];
    const uint32_t* dj = digits + ebase + (size_t)j * d.n;
    uint2* out = coarse + ebase;
    for (uint32_t base = i0; base < i1; base += SC_ROUND) {
        if (tid < MSM_COARSE) cnt[tid] = 0;
        __syncthreads();
        uint32_t val[SC_PER], key[SC_PER], rank[SC_PER];
        for (uint32_t k = 0; k < SC_PER; ++k) {
            const uint32_t i = base + tid + k * SC_THREADS;
            rank[k] = SC_NONE;
            if (i < i1) {
                const uint32_t dg = dj[i], m = dg & ~DIGIT_NEG;
                if (m) {
                    key[k] = m - 1;
                    val[k] = i | (dg & DIGIT_NEG);
                    rank[k] = atomicAdd(&cnt[(m - 1) >> fshift], 1u);
                }
            }
        }
        __syncthreads();
        lds_exclusive_scan(cnt, lstart, tmp, MSM_COARSE);
        for (uint32_t k = 0; k < SC_PER; ++k) {
            if (rank[k] == SC_NONE) continue;
            const uint32_t h = key[k] >> fshift, p = lstart[h] + rank[k];
            stage[p] = make_uint2(val[k], key[k] & (FB - 1));
            skey[p] = (uint8_t)h;
        }
        __syncthreads();
        const uint32_t total = lstart[MSM_COARSE - 1] + cnt[MSM_COARSE - 1];
        if (d.tp == 2) {   // packed: point index < 2^24 and fine key < 128 share one word with the sign
            uint32_t* out32 = (uint32_t*)out;
            for (uint32_t p = tid; p < total; p += SC_THREADS) {
                const uint32_t h = skey[p];
                out32[cur[h] + (p - lstart[h])] = stage[p].x | (stage[p].y << 24);
            }
        } else {
            for (uint32_t p = tid; p < total; p += SC_THREADS) {
                const uint32_t h = skey[p];
                out[cur[h] + (p - lstart[h])] = stage[p];
            }
        }
        __syncthreads();
        if (tid < MSM_COARSE) cur[tid] += cnt[tid];
        __syncthreads();
    }
}

// pass 2: workgroup per (window, coarse group): bucket order inside the group
__global__ void __launch_bounds__(SC_THREADS) k_msm_scatter_fine(MsmBatchDev b, const uint2* __restrict__ coarse,
                                                                 const uint32_t* __restrict__ offs, uint32_t* __restrict__ sorted) {
    __shared__ uint32_t cur[MSM_MAX_FINE], cnt[MSM_MAX_FINE], lstart[MSM_MAX_FINE], tmp[MSM_MAX_FINE];
    __shared__ uint32_t stage[SC_ROUND];
    __shared__ uint8_t skey[SC_ROUND];
    const uint64_t* beg = space(b, SP_CGROUP);
    const uint32_t ii = find_inst(beg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    if (!d.tp) return;
    const uint32_t local = (uint32_t)(blockIdx.x - beg[ii]);
    const uint32_t FB = d.NB / MSM_COARSE;
    const uint32_t* of = offs + space(b, SP_KEY)[ii] + (size_t)local * FB;   // local = j * 256 + h
    const uint32_t tid = threadIdx.x;
    const uint32_t g0 = of[0], g1 = of[FB];
    if (tid < FB) cur[tid] = of[tid];
    const uint64_t ebase = space(b, SP_ENTRY)[ii];
    const uint2* in = coarse + ebase;
    const uint32_t* in32 = (const uint32_t*)in;   // packed entries (tp == 2)
    const bool packed = d.tp == 2;
    uint32_t* out = sorted + ebase;
    for (uint32_t base = g0; base < g1; base += SC_ROUND) {
        if (tid < FB) cnt[tid] = 0;
        __syncthreads();
        uint2 v[SC_PER];
        uint32_t rank[SC_PER];
        for (uint32_t k = 0; k < SC_PER; ++k) {
            const uint32_t e = base + tid + k * SC_THREADS;
            rank[k] = SC_NONE;
            if (e < g1) {
                if (packed) {
                    const uint32_t w = in32[e];
                    v[k] = make_uint2(w & 0x80ffffffu, (w >> 24) & 0x7fu);
                } else {
                    v[k] = in[e];
                }
                rank[k] = atomicAdd(&cnt[v[k].y], 1u);
            }
        }
        __syncthreads();
        lds_exclusive_scan(cnt, lstart, tmp, FB);
        for (uint32_t k = 0; k < SC_PER; ++k) {
            if (rank[k] == SC_NONE) continue;
            const uint32_t p = lstart[v[k].y] + rank[k];
            stage[p] = v[k].x;
            skey[p] = (uint8_t)v[k].y;
        }
        __syncthreads();
        const uint32_t total = min(SC_ROUND, g1 - base);
        for (uint32_t p = tid; p < total; p += SC_THREADS) {
            const uint32_t f = skey[p];
            out[cur[f] + (p - lstart[f])] = stage[p];
        }
        __syncthreads();
        if (tid < FB) cur[tid] += cnt[tid];
        __syncthreads();
    }
}

// ------------------------------------------------------------------ host side
void msm_prepare() {
    static std::once_flag once;
    std::call_once(once, [] {
        int bytes = (1 << (MSM_MAX_C - 1)) * 4;
        (void)hipFuncSetAttribute((const void*)k_msm_hist, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_msm_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    });
}

static unsigned grid_for(uint64_t n, unsigned threads, unsigned cap) {
    uint64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    return (unsigned)(g < cap ? g : cap);
}

void msm_launch_scalar_stages(hzk_ctx* ctx, const MsmBatchDev& b, const uint64_t* tot, uint32_t max_nb,
                              const MsmScalarBufs& bufs) {
    hipStream_t st = ctx->stream;
    const uint64_t ntiles = tot[SP_SCANTILE];
    {
        TimedLaunch tl(ctx, "msm_digits");
        hipLaunchKernelGGL(k_msm_digits, dim3(grid_for(tot[SP_DIGIT], 256, 1u << 16)), dim3(256), 0, st, b, bufs.digits);
    }
    {
        TimedLaunch tl(ctx, "msm_hist");
        hipLaunchKernelGGL(k_msm_hist, dim3((unsigned)tot[SP_TILEWG]), dim3(1024), max_nb * 4, st, b, (const uint32_t*)bufs.digits,
                           bufs.hist);
    }
    {
        TimedLaunch tl(ctx, "msm_scan");
        hipLaunchKernelGGL(k_msm_colscan, dim3(grid_for(tot[SP_KEY], 256, 1u << 30)), dim3(256), 0, st, b, bufs.hist, bufs.count);
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, st, b, (const uint32_t*)bufs.count,
                           bufs.tiles, bufs.tiles + ntiles);
        hipLaunchKernelGGL(k_scan_top, dim3(b.ninst), dim3(1024), 0, st, b, bufs.tiles, (const uint64_t*)(bufs.tiles + ntiles));
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(SCAN_THREADS), 0, st, b, (const uint32_t*)bufs.count,
                           (const uint64_t*)bufs.tiles, bufs.offs);
    }
    {
        TimedLaunch tl(ctx, "msm_scatter");
        hipLaunchKernelGGL(k_msm_scatter, dim3((unsigned)tot[SP_TILEWG]), dim3(1024), max_nb * 4, st, b,
                           (const uint32_t*)bufs.digits, (const uint32_t*)bufs.hist, (const uint32_t*)bufs.offs, bufs.sorted);
        if (bufs.coarse) {
            hipLaunchKernelGGL(k_msm_scatter_coarse, dim3((unsigned)tot[SP_TILEWG]), dim3(SC_THREADS), 0, st, b,
                               (const uint32_t*)bufs.digits, (const uint32_t*)bufs.hist, (const uint32_t*)bufs.offs, bufs.coarse);
            hipLaunchKernelGGL(k_msm_scatter_fine, dim3((unsigned)tot[SP_CGROUP]), dim3(SC_THREADS), 0, st, b,
                               (const uint2*)bufs.coarse, (const uint32_t*)bufs.offs, bufs.sorted);
        }
    }
}

// Built-in plan (autotuned entries override it per size class).
// c minimises n*W (bucket entries, one mixed addition each) + 3*W*2^(c-1)
// (bucket reduction: two full additions per bucket plus its share of the run
// weights), within the LDS histogram limit. K makes the accumulation about
// 2^18..2^19 lanes (4 waves per SIMD on every CU), each adding exactly K
// entries.
MsmPlan msm_plan(size_t n, int force_c, int force_g, size_t lds_bytes) {
    MsmPlan p;
    int cmax = MSM_MAX_C;
    while (cmax > MSM_MIN_C && ((size_t)4 << (cmax - 1)) > lds_bytes) --cmax;   // histogram must fit in LDS
    int c = MSM_MIN_C;
    double best = -1;
    for (int cc = MSM_MIN_C; cc <= cmax; ++cc) {
        double w = (double)((255 + cc - 1) / cc);
        double cost = (double)n * w + 3.0 * w * (double)(1u << (cc - 1));
        if (best < 0 || cost < best) { best = cost; c = cc; }
    }
    if (force_c) c = force_c;
    p.c = c;
    p.W = (255 + c - 1) / c;
    p.NB = 1u << (c - 1);
    size_t L = n * (size_t)p.W;
    uint32_t K = 4;   // multiple of 4: the accumulation reads indices in groups of four
    while (K < 4096 && L / (2 * (size_t)K) >= ((size_t)1 << 18)) K <<= 1;
    p.K = K;
    uint32_t g = p.NB / 512;
    if (g < 4) g = 4;
    if (g > 16) g = 16;
    if (g > p.NB) g = p.NB;
    if (force_g) g = force_g;
    if (g > p.NB) g = p.NB;   // a run never spans more buckets than a window has
    p.g = g;
    // scalar tiles per window for the LDS histogram / scatter workgroups
    // (about 512 workgroups per MSM; every tile at least 4096 scalars and 4x
    // the bucket count, so the per-tile histograms stay smaller than the digits)
    uint32_t T = (512 + p.W - 1) / p.W;
    while (T > 1 && (n / T < 4096 || n / T < 4 * (size_t)p.NB)) T >>= 1;
    p.T = T;
    p.tile = (uint32_t)((n + T - 1) / T);
    return p;
}

// plans measured by the autotuner on an MI355X (liblightzk_amd/tune/mi355x.json),
// [regime][G1, G2][log2 n]
static const hzk_tune_entry BUILTIN_PLANS[2][2][32] = {
#include "tune_table.inc"
};

static int size_class(size_t n) {
    int cls = 0;
    while (((size_t)1 << (cls + 1)) <= n) ++cls;
    return cls < 32 ? cls : 31;
}

static bool plan_from_entry(const hzk_tune_entry& t, size_t n, int force_g, size_t lds_bytes, MsmPlan* out) {
    if (!t.c || ((size_t)4 << (t.c - 1)) > lds_bytes) return false;
    MsmPlan q = msm_plan(n, t.c, t.g ? (int)t.g : force_g, lds_bytes);
    if (t.K) q.K = t.K;
    *out = q;
    return true;
}

MsmPlan msm_plan_builtin(int g2, size_t n, int regime, size_t lds_bytes) {
    MsmPlan p = msm_plan(n, 0, 0, lds_bytes);
    const int cls = size_class(n);
    if (regime == MSM_REGIME_BATCH && plan_from_entry(BUILTIN_PLANS[1][g2 ? 1 : 0][cls], n, 0, lds_bytes, &p)) return p;
    plan_from_entry(BUILTIN_PLANS[0][g2 ? 1 : 0][cls], n, 0, lds_bytes, &p);
    return p;
}

// plan for an MSM of n points: the context's tuned entry for the regime and
// size class, else the built-in measured plan of the regime; a batch-regime
// class with neither falls back to the single-MSM plan, and that one to the
// heuristic. A caller-forced window size (or hzk_set_tuning) bypasses the
// tables.
MsmPlan msm_plan_ctx(hzk_ctx* ctx, int g2, size_t n, int force_c, int regime) {
    MsmPlan p = msm_plan(n, force_c ? force_c : ctx->force_window, ctx->force_seg, ctx->lds_per_block);
    if (!force_c && !ctx->force_window) {
        const int cls = size_class(n), gi = g2 ? 1 : 0;
        for (int r = regime; r >= MSM_REGIME_SINGLE; --r) {
            hzk_tune_entry t = ctx->tune[r][gi][cls];
            if (!t.c && !ctx->force_seg) t = BUILTIN_PLANS[r][gi][cls];
            if (plan_from_entry(t, n, ctx->force_seg, ctx->lds_per_block, &p)) break;
        }
    }
    return p;
}

}  // namespace hzk
