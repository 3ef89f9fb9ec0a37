// Group-generic MSM stages 5-8 and the batched driver (G1 over Fq, G2 over
// Fq2); included by msm_g1.hip and msm_g2.hip, one instantiation per
// translation unit. Stages 1-4 and the batch layout are described in msm.hip.
#pragma once
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>
#include "common.hpp"
#include "curve.hpp"
#include "field29.hpp"
#include "host_ec.hpp"
#include "host_pool.hpp"
#include "msm.hpp"

namespace hzk {

// fan-in of the tree over a hot bucket's chunk partials: each level is a
// chain of K2 - 1 dependent additions on few lanes, so latency, not work,
// sets its cost (4: three additions per level)
static constexpr int K2 = 4;
static constexpr uint32_t HOT = 16;  // a bucket spanning more chunks than this uses the tree
static constexpr uint32_t HOT_LIST = 256;   // hot instances a hot-tree block lists in LDS

HZK_DEV uint32_t find_inst_g(const uint64_t* __restrict__ beg, uint32_t ninst, uint64_t gid) {
    uint32_t lo = 0, hi = ninst;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (beg[mid] <= gid) lo = mid; else hi = mid;
    }
    return lo;
}
HZK_DEV const uint64_t* space_g(const MsmBatchDev& b, int s) { return b.begin + (size_t)s * (b.ninst + 1); }

// the key whose entries [offs[k], offs[k+1]) contain entry e (offs[nk] > e)
HZK_DEV uint32_t find_key(const uint32_t* __restrict__ offs, uint32_t nk, uint32_t e) {
    uint32_t lo = 0, hi = nk;   // invariant: offs[lo] <= e < offs[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------ 5. accumulate
// Lane per chunk of K consecutive sorted entries; a chunk crosses bucket
// boundaries freely, so every lane does exactly K mixed additions (the last
// chunk of an instance fewer). The running XYZZ sum is flushed at each bucket
// end: a bucket that lies wholly inside the chunk is final and goes straight
// to bucket[k]; the partial of a bucket that began in an earlier chunk goes
// to carry[2c] (head), that of a bucket continuing into the next chunk to
// carry[2c+1] (tail). Persistent grid over the whole batch's chunk space.
// G1: the additions in 29-bit-limb arithmetic (field29.hpp: one
// v_mad_u64_u32 per partial product, lazy reduction) over the bases converted
// to its Montgomery form by k_msm_convert_g1. Bucket sums and chunk partials
// are stored as raw accumulator states (g1acc29::store_raw), the form the
// fix-up and the bucket reduction work in. At least two waves per SIMD (the
// G2 lane state does not fit twice in 512 registers otherwise).
static __global__ void __launch_bounds__(256, 2) k_msm_accumulate_g1(MsmBatchDev b, const uint32_t* __restrict__ sorted,
                                                           const uint32_t* __restrict__ offs, uint4* __restrict__ bucket,
                                                           uint4* __restrict__ carry, uint32_t* __restrict__ chunk2key) {
    constexpr int RW = g1acc29::RAW_WORDS / 4;   // uint4 per raw point
    const uint64_t* cbeg = space_g(b, SP_CHUNK);
    const uint64_t total = cbeg[b.ninst];
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t ii = find_inst_g(cbeg, b.ninst, gid);
        const MsmInst& d = b.inst[ii];
        const uint32_t E = b.stats[4 * ii];
        const uint32_t e0 = (uint32_t)(gid - cbeg[ii]) * d.K;
        if (e0 >= E) continue;
        const uint32_t e1 = min(e0 + d.K, E);
        const uint64_t kb = space_g(b, SP_KEY)[ii];
        const uint32_t* of = offs + kb;
        uint32_t k = find_key(of, d.nk, e0);
        chunk2key[gid] = k;
        uint32_t next = of[k + 1];
        bool before = of[k] < e0;
        const uint4* srt4 = (const uint4*)(sorted + space_g(b, SP_ENTRY)[ii]);
        const uint4* bases = (const uint4*)d.bases;   // converted: x | y, 8 x 32-bit limbs each
        g1acc29 acc;
        acc.set_inf();
        uint4 q = srt4[e0 >> 2];
        uint32_t idx = q.x;
        uint4 p0, p1, p2, p3;
        {
            const uint4* bp = bases + 4 * (size_t)(idx & 0x7fffffffu);
            p0 = bp[0], p1 = bp[1], p2 = bp[2], p3 = bp[3];
        }
        for (uint32_t e = e0; e < e1; ++e) {
            const bool neg = idx >> 31;
            const uint4 x0 = p0, x1 = p1, y0 = p2, y1 = p3;
            const bool inf = p1.w == 0xffffffffu;   // the converter's mark for the point at infinity
            if (e + 1 < e1) {   // prefetch the next base while this one is added
                const uint32_t t = (e + 1) & 3;
                if (t == 0) q = srt4[(e + 1) >> 2];
                idx = t == 0 ? q.x : (t == 1 ? q.y : (t == 2 ? q.z : q.w));
                const uint4* bp = bases + 4 * (size_t)(idx & 0x7fffffffu);
                p0 = bp[0], p1 = bp[1], p2 = bp[2], p3 = bp[3];
            }
            if (!inf) {
                const fq29 x = fq29::unpack(x0, x1), ya = fq29::unpack(y0, y1);
                const fq29 y = neg ? fq29::neg(ya, fq29::P1) : ya;
                acc.add_affine(x, y);
            }
            if (e + 1 == next || e + 1 == e1) {
                if (before) acc.store_raw(carry + RW * (2 * gid));
                else if (next > e1) acc.store_raw(carry + RW * (2 * gid + 1));
                else acc.store_raw(bucket + RW * (kb + k));
                if (e + 1 < e1) {   // next non-empty bucket
                    do { ++k; next = of[k + 1]; } while (next <= e + 1);
                    before = false;
                    acc.set_inf();
                }
            }
        }
    }
}

// G2: the same walk over Fq2 = Fq[u]/(u^2 + 1) in 29-bit-limb components
// (g2acc29), bases converted by k_msm_convert (128 bytes: x.c0 | x.c1 |
// y.c0 | y.c1)
static __global__ void __launch_bounds__(256, 2) k_msm_accumulate_g2(MsmBatchDev b, const uint32_t* __restrict__ sorted,
                                                           const uint32_t* __restrict__ offs, uint4* __restrict__ bucket,
                                                           uint4* __restrict__ carry, uint32_t* __restrict__ chunk2key) {
    constexpr int RW = g2acc29::RAW_WORDS / 4;
    const uint64_t* cbeg = space_g(b, SP_CHUNK);
    const uint64_t total = cbeg[b.ninst];
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t ii = find_inst_g(cbeg, b.ninst, gid);
        const MsmInst& d = b.inst[ii];
        const uint32_t E = b.stats[4 * ii];
        const uint32_t e0 = (uint32_t)(gid - cbeg[ii]) * d.K;
        if (e0 >= E) continue;
        const uint32_t e1 = min(e0 + d.K, E);
        const uint64_t kb = space_g(b, SP_KEY)[ii];
        const uint32_t* of = offs + kb;
        uint32_t k = find_key(of, d.nk, e0);
        chunk2key[gid] = k;
        uint32_t next = of[k + 1];
        bool before = of[k] < e0;
        const uint4* srt4 = (const uint4*)(sorted + space_g(b, SP_ENTRY)[ii]);
        const uint4* bases = (const uint4*)d.bases;
        g2acc29 acc;
        acc.set_inf();
        uint4 q = srt4[e0 >> 2];
        uint32_t idx = q.x;
        for (uint32_t e = e0; e < e1; ++e) {
            const bool neg = idx >> 31;
            const uint4* bp = bases + 8 * (size_t)(idx & 0x7fffffffu);
            const uint4 p0 = bp[0], p1 = bp[1], p2 = bp[2], p3 = bp[3], p4 = bp[4], p5 = bp[5], p6 = bp[6], p7 = bp[7];
            if (e + 1 < e1) {
                const uint32_t t = (e + 1) & 3;
                if (t == 0) q = srt4[(e + 1) >> 2];
                idx = t == 0 ? q.x : (t == 1 ? q.y : (t == 2 ? q.z : q.w));
            }
            if (p3.w != 0xffffffffu) {   // the converter's mark for the point at infinity
                const fq2_29 x{fq29::unpack(p0, p1), fq29::unpack(p2, p3)};
                fq2_29 y{fq29::unpack(p4, p5), fq29::unpack(p6, p7)};
                if (neg) y = {fq29::neg(y.c0, fq29::P1), fq29::neg(y.c1, fq29::P1)};
                acc.add_affine(x, y);
            }
            if (e + 1 == next || e + 1 == e1) {
                if (before) acc.store_raw(carry + RW * (2 * gid));
                else if (next > e1) acc.store_raw(carry + RW * (2 * gid + 1));
                else acc.store_raw(bucket + RW * (kb + k));
                if (e + 1 < e1) {
                    do { ++k; next = of[k + 1]; } while (next <= e + 1);
                    before = false;
                    acc.set_inf();
                }
            }
        }
    }
}

// one base array to convert for the G1 accumulation: count affine points
// from src into dst, as items [begin, next begin) of the conversion space
struct MsmConvTask {
    const void* src;
    void* dst;
    uint64_t begin;
};

// x, y -> x 2^5, y 2^5 mod p (the 2^261 Montgomery form of field29.hpp),
// canonical, 8 x 32-bit limbs; the point at infinity (all zero) becomes
// x word 7 = 0xffffffff, a value no canonical coordinate has
HZK_DEV void convert_point_g1(const uint4* __restrict__ sp, uint4* __restrict__ dp) {
    const uint4 a = sp[0], b2 = sp[1], c = sp[2], e = sp[3];
    const uint32_t any = a.x | a.y | a.z | a.w | b2.x | b2.y | b2.z | b2.w | c.x | c.y | c.z | c.w | e.x | e.y | e.z | e.w;
    if (!any) {
        dp[0] = make_uint4(0, 0, 0, 0);
        dp[1] = make_uint4(0, 0, 0, 0xffffffffu);
        dp[2] = make_uint4(0, 0, 0, 0);
        dp[3] = make_uint4(0, 0, 0, 0);
        return;
    }
    const uint32_t xw[8] = {a.x, a.y, a.z, a.w, b2.x, b2.y, b2.z, b2.w};
    const uint32_t yw[8] = {c.x, c.y, c.z, c.w, e.x, e.y, e.z, e.w};
    uint32_t xo[8], yo[8];
    fq29 x, y;
    fq29::mul2(x, fq29::unpack(xw), fq29::konst(fq29::TO), y, fq29::unpack(yw), fq29::konst(fq29::TO));
    x.canonical().pack(xo);
    y.canonical().pack(yo);
    dp[0] = make_uint4(xo[0], xo[1], xo[2], xo[3]);
    dp[1] = make_uint4(xo[4], xo[5], xo[6], xo[7]);
    dp[2] = make_uint4(yo[0], yo[1], yo[2], yo[3]);
    dp[3] = make_uint4(yo[4], yo[5], yo[6], yo[7]);
}

// G2: x.c0 | x.c1 | y.c0 | y.c1 -> the same conversion per Fq component;
// infinity (all zero) marked by x.c1 word 7 = 0xffffffff
HZK_DEV void convert_point_g2(const uint4* __restrict__ sp, uint4* __restrict__ dp) {
    uint4 v[8];
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = sp[i];
        any |= v[i].x | v[i].y | v[i].z | v[i].w;
    }
    if (!any) {
#pragma unroll
        for (int i = 0; i < 8; ++i) dp[i] = make_uint4(0, 0, 0, i == 3 ? 0xffffffffu : 0);
        return;
    }
#pragma unroll
    for (int c = 0; c < 4; c += 2) {
        const uint32_t a[8] = {v[2 * c].x, v[2 * c].y, v[2 * c].z, v[2 * c].w, v[2 * c + 1].x, v[2 * c + 1].y, v[2 * c + 1].z, v[2 * c + 1].w};
        const uint32_t b[8] = {v[2 * c + 2].x, v[2 * c + 2].y, v[2 * c + 2].z, v[2 * c + 2].w,
                               v[2 * c + 3].x, v[2 * c + 3].y, v[2 * c + 3].z, v[2 * c + 3].w};
        fq29 x, y;
        fq29::mul2(x, fq29::unpack(a), fq29::konst(fq29::TO), y, fq29::unpack(b), fq29::konst(fq29::TO));
        uint32_t xo[8], yo[8];
        x.canonical().pack(xo);
        y.canonical().pack(yo);
        dp[2 * c] = make_uint4(xo[0], xo[1], xo[2], xo[3]);
        dp[2 * c + 1] = make_uint4(xo[4], xo[5], xo[6], xo[7]);
        dp[2 * c + 2] = make_uint4(yo[0], yo[1], yo[2], yo[3]);
        dp[2 * c + 3] = make_uint4(yo[4], yo[5], yo[6], yo[7]);
    }
}

static __global__ void __launch_bounds__(256) k_msm_convert_g2(const MsmConvTask* __restrict__ tasks, uint32_t ntasks, uint64_t total) {
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ntasks;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tasks[mid].begin <= gid) lo = mid; else hi = mid;
        }
        const MsmConvTask& t = tasks[lo];
        const uint64_t i = gid - t.begin;
        convert_point_g2((const uint4*)t.src + 8 * i, (uint4*)t.dst + 8 * i);
    }
}

static __global__ void __launch_bounds__(256) k_msm_convert_g1(const MsmConvTask* __restrict__ tasks, uint32_t ntasks, uint64_t total) {
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ntasks;   // tasks[lo].begin <= gid < tasks[hi].begin
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tasks[mid].begin <= gid) lo = mid; else hi = mid;
        }
        const MsmConvTask& t = tasks[lo];
        const uint64_t i = gid - t.begin;
        convert_point_g1((const uint4*)t.src + 4 * i, (uint4*)t.dst + 4 * i);
    }
}

// one array, in place or not
static __global__ void __launch_bounds__(256) k_msm_convert_g1_array(const uint4* src, uint4* dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        convert_point_g1(src + 4 * i, dst + 4 * i);
}
static __global__ void __launch_bounds__(256) k_msm_convert_g2_array(const uint4* src, uint4* dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        convert_point_g2(src + 8 * i, dst + 8 * i);
}

// ------------------------------------------------------------------ 6. fix-up
// The 29-bit accumulator of each group; bucket sums, chunk partials and run
// sums live in memory as its raw states (RW uint4 per point).
template <class F> struct Acc29;
template <> struct Acc29<fq_t> { using A = g1acc29; };
template <> struct Acc29<fq2_t> { using A = g2acc29; };
template <class F> using acc29_t = typename Acc29<F>::A;
template <class F> constexpr int raw_vec() { return acc29_t<F>::RAW_WORDS / 4; }

// hot buckets (more than HOT chunks): one level of an in-place K2-ary tree
// over the head partials carry[2c] of chunks f+1..l (f, l: first and last
// chunk of the bucket); afterwards carry[2(f+1)] holds their sum
template <class F>
HZK_DEV void hot_tree_chunk(const MsmBatchDev& b, const uint32_t* __restrict__ offs, const uint32_t* __restrict__ chunk2key,
                            uint32_t stride, uint4* __restrict__ carry, uint32_t ii, uint32_t c) {
    using A = acc29_t<F>;
    constexpr int RW = raw_vec<F>();
    const MsmInst& d = b.inst[ii];
    if ((uint64_t)c * d.K >= b.stats[4 * ii]) return;
    const uint64_t* cbeg = space_g(b, SP_CHUNK);
    const uint32_t* of = offs + space_g(b, SP_KEY)[ii];
    const uint32_t k = chunk2key[cbeg[ii] + c], start = of[k];
    if (start >= c * d.K) return;   // chunk c holds no head partial
    const uint32_t f = start / d.K, l = (of[k + 1] - 1) / d.K;
    if (l - f <= HOT) return;
    const uint32_t rel = c - (f + 1), span = l - f;
    if (rel % (stride * K2) != 0) return;
    uint4* base = carry + (size_t)RW * 2 * cbeg[ii];
    A acc = A::load_raw(base + (size_t)RW * 2 * c);
    for (int i = 1; i < K2; ++i) {
        uint32_t r = rel + i * stride;
        if (r >= span) break;
        acc.add(A::load_raw(base + (size_t)RW * 2 * (c + i * stride)));
    }
    acc.store_raw(base + (size_t)RW * 2 * c);
}

template <class F>
__global__ void __launch_bounds__(256) k_msm_hot_tree(MsmBatchDev b, const uint32_t* __restrict__ offs,
                                                      const uint32_t* __restrict__ chunk2key, uint32_t stride,
                                                      uint4* __restrict__ carry) {
    // Most instances have no hot bucket. Each block lists the instances that
    // do (largest bucket > HOT chunks), in instance order, with the prefix
    // sums of their chunk counts, all in LDS; the grid then strides over the
    // concatenated chunks of the hot instances only, one (instance, chunk)
    // per lane step.
    __shared__ uint32_t hot[HOT_LIST];
    __shared__ uint64_t hbeg[HOT_LIST + 1];
    __shared__ uint64_t wmask[4];
    __shared__ uint32_t nhot;
    const uint64_t* cbeg = space_g(b, SP_CHUNK);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) nhot = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < b.ninst; i0 += 256) {
        const uint32_t ii = i0 + threadIdx.x;
        const bool h = ii < b.ninst && b.stats[4 * ii + 2] > HOT * b.inst[ii].K;
        const uint64_t m = __ballot(h);
        if (lane == 0) wmask[wave] = m;
        __syncthreads();
        uint32_t pos = nhot + __popcll(m & ((1ull << lane) - 1));
        for (uint32_t w = 0; w < wave; ++w) pos += __popcll(wmask[w]);
        if (h && pos < HOT_LIST) hot[pos] = ii;
        __syncthreads();
        if (threadIdx.x == 0) nhot += __popcll(wmask[0]) + __popcll(wmask[1]) + __popcll(wmask[2]) + __popcll(wmask[3]);
        __syncthreads();
    }
    const uint32_t nh = nhot;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    if (nh > HOT_LIST) {   // more hot instances than the list holds: walk every chunk
        for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < cbeg[b.ninst]; gid += step) {
            const uint32_t ii = find_inst_g(cbeg, b.ninst, gid);
            if (b.stats[4 * ii + 2] > HOT * b.inst[ii].K)
                hot_tree_chunk<F>(b, offs, chunk2key, stride, carry, ii, (uint32_t)(gid - cbeg[ii]));
        }
        return;
    }
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (uint32_t h = 0; h < nh; ++h) { hbeg[h] = t; t += cbeg[hot[h] + 1] - cbeg[hot[h]]; }
        hbeg[nh] = t;
    }
    __syncthreads();
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < hbeg[nh]; g += step) {
        uint32_t lo = 0, hi = nh;   // hbeg[lo] <= g < hbeg[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (hbeg[mid] <= g) lo = mid; else hi = mid;
        }
        hot_tree_chunk<F>(b, offs, chunk2key, stride, carry, hot[lo], (uint32_t)(g - hbeg[lo]));
    }
}

// bucket spanning chunks f..l (f < l): bucket = tail(f) + heads(f+1..l), by
// the lane of chunk f+1 (the first chunk that continues it)
template <class F>
__global__ void __launch_bounds__(256) k_msm_fixup(MsmBatchDev b, const uint32_t* __restrict__ offs,
                                                   const uint32_t* __restrict__ chunk2key, const uint4* __restrict__ carry,
                                                   uint4* __restrict__ bucket) {
    using A = acc29_t<F>;
    constexpr int RW = raw_vec<F>();
    const uint64_t* cbeg = space_g(b, SP_CHUNK);
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= cbeg[b.ninst]) return;
    uint32_t ii = find_inst_g(cbeg, b.ninst, gid);
    const MsmInst& d = b.inst[ii];
    const uint32_t c = (uint32_t)(gid - cbeg[ii]);
    if ((uint64_t)c * d.K >= b.stats[4 * ii]) return;
    const uint64_t kb = space_g(b, SP_KEY)[ii];
    const uint32_t k = chunk2key[gid], start = offs[kb + k];
    if (start >= c * d.K) return;   // the chunk's first bucket starts in it
    const uint32_t f = start / d.K;
    if (c != f + 1) return;
    const uint32_t l = (offs[kb + k + 1] - 1) / d.K;
    const uint4* base = carry + (size_t)RW * 2 * cbeg[ii];
    A acc = A::load_raw(base + (size_t)RW * (2 * f + 1));
    if (l - f > HOT) {
        acc.add(A::load_raw(base + (size_t)RW * 2 * (f + 1)));
    } else {
        for (uint32_t cc = f + 1; cc <= l; ++cc) acc.add(A::load_raw(base + (size_t)RW * 2 * cc));
    }
    acc.store_raw(bucket + (size_t)RW * (kb + k));
}

// ------------------------------------------------------------------ 7. bucket reduction
// A window's value is sum_{d=1..NB} d * B_d. Bucket d = r*g + q + 1 lies in
// run r (g buckets) at offset q, so the sum is
//     sum_r [ A_r + g * r * R_r ],   A_r = sum_q (q+1) B,   R_r = sum_q B.
// Stage 7a: lane per (window, run): A_r and R_r by running sums (2g
// additions, no scalar multiplication). Stage 7b: workgroup per (window,
// part of Gs = G/S runs): thread t takes m = Gs/T consecutive runs; the run
// weights r become t*m + i inside the part, and sum_t t * rr_t (rr_t the
// thread's run sum) is the sum over t >= 1 of the suffix sums of rr, one
// LDS scan. Each part yields its value P_p (weights relative to the part's
// first run) and its run sum RS_p; the host adds g * p * Gs * RS_p. Buckets,
// run sums and the LDS slots hold raw accumulator states; only P_p and RS_p
// are converted to R = 2^256 XYZZ points for the host.
template <class F>
__global__ void __launch_bounds__(128) k_bucket_segments(MsmBatchDev b, const uint4* __restrict__ bucket,
                                                         const uint32_t* __restrict__ offs, uint4* __restrict__ segsum) {
    using A = acc29_t<F>;
    constexpr int RW = raw_vec<F>();
    const uint64_t* rbeg = space_g(b, SP_BRED);
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= rbeg[b.ninst]) return;
    uint32_t ii = find_inst_g(rbeg, b.ninst, gid);
    const MsmInst& d = b.inst[ii];
    if (d.fused) return;   // its window-sum kernel reduces the runs itself
    uint32_t t = (uint32_t)(gid - rbeg[ii]);
    uint32_t j = t / d.G, sg = t % d.G;
    const uint64_t kb = space_g(b, SP_KEY)[ii];
    const uint32_t* of = offs + kb;
    const uint4* bk = bucket + (size_t)RW * kb;
    uint32_t kbase = j * d.NB + sg * d.g;
    A run, acc;
    run.set_inf();
    acc.set_inf();
    for (int q = (int)d.g - 1; q >= 0; --q) {
        uint32_t key = kbase + q;
        if (of[key + 1] != of[key]) run.add(A::load_raw(bk + (size_t)RW * key));
        acc.add(run);
    }
    acc.store_raw(segsum + (size_t)RW * 2 * gid);         // A_r
    run.store_raw(segsum + (size_t)RW * (2 * gid + 1));   // R_r
}

// Per-thread accumulator slots in LDS, word-major: word w of slot t at
// sh[w * T + t], so the lanes of a wave touch consecutive words (a point per
// slot in a row would put every lane's word w in the same bank)
template <class F, int T>
struct LdsPoints {
    using A = acc29_t<F>;
    static constexpr int W = A::RAW_WORDS;
    uint32_t* sh;
    HZK_DEV void put(uint32_t t, const A& p) const {
        uint32_t w[W];
        p.to_words(w);
#pragma unroll
        for (int i = 0; i < W; ++i) sh[i * T + t] = w[i];
    }
    HZK_DEV A get(uint32_t t) const {
        uint32_t w[W];
#pragma unroll
        for (int i = 0; i < W; ++i) w[i] = sh[i * T + t];
        return A::from_words(w);
    }
};

// tree sum of slots 0..T) into slot 0
template <class F, int T>
HZK_DEV void lds_tree_sum(const LdsPoints<F, T>& sh, uint32_t t) {
    for (uint32_t off = T / 2; off > 0; off >>= 1) {
        if (t < off) {
            auto x = sh.get(t);
            x.add(sh.get(t + off));
            sh.put(t, x);
        }
        __syncthreads();
    }
}

// T threads per part: 256 for single MSMs (shallow: the kernel is latency
// bound there), 64 in wide batches (each thread takes 4x more runs, so the
// scan and tree overhead per run is 4x smaller). Dynamic LDS: RAW_WORDS * T words.
template <class F, int T>
__global__ void __launch_bounds__(T) k_window_sum(MsmBatchDev b, const uint4* __restrict__ segsum,
                                                  const uint4* __restrict__ bucket, const uint32_t* __restrict__ offs,
                                                  xyzz_t<F>* __restrict__ winsum) {
    using A = acc29_t<F>;
    constexpr int RW = raw_vec<F>();
    extern __shared__ uint32_t shw[];
    const LdsPoints<F, T> sh{shw};
    const uint64_t* wbeg = space_g(b, SP_WIN);
    uint32_t ii = find_inst_g(wbeg, b.ninst, blockIdx.x);
    const MsmInst& d = b.inst[ii];
    const uint32_t local = (uint32_t)(blockIdx.x - wbeg[ii]);
    const uint32_t j = local / d.S, part = local % d.S, Gs = d.G / d.S;
    const uint32_t m = Gs >= T ? Gs / T : 1, t = threadIdx.x;
    // runs t*m .. t*m + m - 1 of the part (A_r at 2r, R_r at 2r + 1)
    const uint4* ss =
        segsum + (size_t)RW * 2 * (space_g(b, SP_BRED)[ii] + (size_t)j * d.G + (size_t)part * Gs + (size_t)t * m);
    A a, run, w;
    a.set_inf();
    run.set_inf();
    w.set_inf();
    if (d.fused) {
        // m = 1: this lane's run r = part * T + t, buckets r*g .. r*g + g-1 of
        // window j; A_r and R_r by running sums as in k_bucket_segments
        const uint64_t kb = space_g(b, SP_KEY)[ii];
        const uint32_t* of = offs + kb;
        const uint4* bk = bucket + (size_t)RW * kb;
        const uint32_t kbase = j * d.NB + (part * T + t) * d.g;
        for (int q = (int)d.g - 1; q >= 0; --q) {
            const uint32_t key = kbase + q;
            if (of[key + 1] != of[key]) run.add(A::load_raw(bk + (size_t)RW * key));
            a.add(run);
        }
    } else if (t * m < Gs) {
        for (int i = (int)m - 1; i >= 0; --i) {
            a.add(A::load_raw(ss + (size_t)RW * 2 * i));
            run.add(A::load_raw(ss + (size_t)RW * (2 * i + 1)));
            if (i > 0) w.add(run);   // w = sum_i i * R_i
        }
    }
    // inclusive suffix scan of the thread run sums: slot t = sum_{t' >= t} rr_t'
    sh.put(t, run);
    __syncthreads();
    for (uint32_t off = 1; off < T; off <<= 1) {
        A x = sh.get(t);
        if (t + off < T) x.add(sh.get(t + off));
        __syncthreads();
        sh.put(t, x);
        __syncthreads();
    }
    const A part_sum = sh.get(0);
    // P = sum_t a_t + g * sum_t (w_t + m * X_t), X_t = slot t for t >= 1
    A u = sh.get(t);
    if (t == 0) u.set_inf();
    for (uint32_t k = 1; k < m; k <<= 1) u.dbl();
    u.add(w);
    if (T == 256) {
        // shallow: the factor g per thread, one tree
        for (uint32_t k = 1; k < d.g; k <<= 1) u.dbl();
        u.add(a);
        __syncthreads();
        sh.put(t, u);
        __syncthreads();
        lds_tree_sum<F, T>(sh, t);
        if (t == 0) {
            winsum[2 * blockIdx.x] = sh.get(0).to_xyzz();   // P_p
            winsum[2 * blockIdx.x + 1] = part_sum.to_xyzz();  // RS_p
        }
    } else {
        // less work: both sums first, the factor g once
        __syncthreads();
        sh.put(t, u);
        __syncthreads();
        lds_tree_sum<F, T>(sh, t);
        A total = sh.get(0);
        __syncthreads();
        sh.put(t, a);
        __syncthreads();
        lds_tree_sum<F, T>(sh, t);
        if (t == 0) {
            for (uint32_t k = 1; k < d.g; k <<= 1) total.dbl();
            total.add(sh.get(0));
            winsum[2 * blockIdx.x] = total.to_xyzz();         // P_p
            winsum[2 * blockIdx.x + 1] = part_sum.to_xyzz();  // RS_p
        }
    }
}

// ------------------------------------------------------------------ 8. device-side combine (optional)
// The host's combine as two kernels, in the 29-bit accumulator arithmetic.
// k_msm_combine_windows: workgroup per (instance, window), lane per part p:
// T_j = sum_p P_p + (g Gs) sum_p p RS_p, the last sum as the sum over p >= 1
// of the suffix sums of RS (an LDS scan), stored as a raw state in the
// run-sum area (free once the window sums exist). k_msm_combine_horner: lane
// per instance, Horner over the windows (the dependent doublings that make
// this path slower than the host's), then the normalised Jacobian
// (x, y, 1) or (0, 1, 0): the bytes the host writes.
template <class F>
HZK_DEV acc29_t<F> acc_from_xyzz(const xyzz_t<F>& p);
template <>
HZK_DEV g1acc29 acc_from_xyzz<fq_t>(const xyzz_t<fq_t>& p) {
    g1acc29 a;
    a.inf = p.ZZ.is_zero();
    fq29::mul2(a.X, fq29::unpack(p.X.l), fq29::konst(fq29::TO), a.Y, fq29::unpack(p.Y.l), fq29::konst(fq29::TO));
    fq29::mul2(a.ZZ, fq29::unpack(p.ZZ.l), fq29::konst(fq29::TO), a.ZZZ, fq29::unpack(p.ZZZ.l), fq29::konst(fq29::TO));
    return a;
}
template <>
HZK_DEV g2acc29 acc_from_xyzz<fq2_t>(const xyzz_t<fq2_t>& p) {
    g2acc29 a;
    a.inf = p.ZZ.is_zero();
    const fq29 T = fq29::konst(fq29::TO);
    fq29::mul2(a.X.c0, fq29::unpack(p.X.c0.l), T, a.X.c1, fq29::unpack(p.X.c1.l), T);
    fq29::mul2(a.Y.c0, fq29::unpack(p.Y.c0.l), T, a.Y.c1, fq29::unpack(p.Y.c1.l), T);
    fq29::mul2(a.ZZ.c0, fq29::unpack(p.ZZ.c0.l), T, a.ZZ.c1, fq29::unpack(p.ZZ.c1.l), T);
    fq29::mul2(a.ZZZ.c0, fq29::unpack(p.ZZZ.c0.l), T, a.ZZZ.c1, fq29::unpack(p.ZZZ.c1.l), T);
    return a;
}

template <class F>
__global__ void __launch_bounds__(64) k_msm_combine_windows(MsmBatchDev b, const xyzz_t<F>* __restrict__ win, uint32_t maxw,
                                                            uint4* __restrict__ tsum) {
    using A = acc29_t<F>;
    constexpr int RW = raw_vec<F>();
    __shared__ uint32_t shw[LdsPoints<F, 64>::W * 64];
    const LdsPoints<F, 64> sh{shw};
    const uint32_t i = blockIdx.x / maxw, j = blockIdx.x % maxw, t = threadIdx.x;
    const MsmInst& d = b.inst[i];
    if (j >= d.W) return;   // uniform per workgroup
    const xyzz_t<F>* pj = win + 2 * (space_g(b, SP_WIN)[i] + (size_t)j * d.S);
    A P, RS;
    P.set_inf();
    RS.set_inf();
    if (t < d.S) {
        P = acc_from_xyzz<F>(pj[2 * t]);
        RS = acc_from_xyzz<F>(pj[2 * t + 1]);
    }
    // inclusive suffix scan of RS: slot t = sum_{p >= t} RS_p
    sh.put(t, RS);
    __syncthreads();
    for (uint32_t off = 1; off < 64; off <<= 1) {
        A x = sh.get(t);
        if (t + off < 64) x.add(sh.get(t + off));
        __syncthreads();
        sh.put(t, x);
        __syncthreads();
    }
    A y = sh.get(t);
    if (t == 0) y.set_inf();   // sum_p p RS_p = sum_{t >= 1} slot t
    __syncthreads();
    sh.put(t, y);
    __syncthreads();
    lds_tree_sum<F, 64>(sh, t);
    A ysum = sh.get(0);
    __syncthreads();
    sh.put(t, P);
    __syncthreads();
    lds_tree_sum<F, 64>(sh, t);
    if (t == 0) {
        A T = sh.get(0);
        for (uint32_t k = 1; k < d.G / d.S * d.g; k <<= 1) ysum.dbl();
        T.add(ysum);
        T.store_raw(tsum + (size_t)RW * (space_g(b, SP_BRED)[i] + j));
    }
}

template <class F>
__global__ void __launch_bounds__(64) k_msm_combine_horner(MsmBatchDev b, const uint4* __restrict__ tsum) {
    using A = acc29_t<F>;
    constexpr int RW = raw_vec<F>();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.ninst) return;
    const MsmInst& d = b.inst[i];
    const uint4* T = tsum + (size_t)RW * space_g(b, SP_BRED)[i];
    A acc;
    acc.set_inf();
    for (int j = (int)d.W - 1; j >= 0; --j) {
        const uint32_t shift = d.Lg > 1 ? d.c * d.Lg : d.wb + ((uint32_t)j < d.wx ? 1u : 0u);
        for (uint32_t q = 0; q < shift; ++q) acc.dbl();
        acc.add(A::load_raw(T + (size_t)RW * j));
    }
    const xyzz_t<F> r = acc.to_xyzz();
    F* out = (F*)d.dest;
    if (r.is_inf()) {
        out[0] = F::zero(), out[1] = F::one(), out[2] = F::zero();
        return;
    }
    const F inv = (r.ZZ * r.ZZZ).inverse();   // 1 / (ZZ ZZZ)
    out[0] = r.X * (r.ZZZ * inv);             // X / ZZ
    out[1] = r.Y * (r.ZZ * inv);              // Y / ZZZ
    out[2] = F::one();
}

// ------------------------------------------------------------------ driver
struct MsmBatchLayout {
    std::vector<MsmInst> inst;
    std::vector<uint64_t> begin;   // [SP_COUNT][ninst + 1]
    uint64_t tot[SP_COUNT];
    uint32_t max_nb = 1;
    uint32_t tree_levels = 0;      // tree levels for hot buckets (upper bound)
    bool two_pass = false;         // some instance scatters in two passes
    int regime = 0;                // plan regime (MSM_REGIME_*)
    size_t bytes = 0;              // scratch needed
    // scratch offsets
    size_t o_inst, o_begin, o_stats, o_digits, o_hist, o_count, o_offs, o_tiles, o_sorted, o_coarse, o_chunk2key, o_bucket,
        o_carry, o_segsum, o_win;
    // base arrays converted for the accumulation, one per distinct source
    // array (its largest referenced point count), and which one each instance reads
    std::vector<MsmConvTask> conv;     // dst holds the offset into the scratch until run time
    std::vector<uint32_t> conv_of;     // instance -> task
    uint64_t conv_total = 0;
    size_t o_conv = 0, o_convdesc = 0;
};

template <class F>
static MsmBatchLayout msm_layout(hzk_ctx* ctx, const MsmJob* jobs, size_t count) {
    MsmBatchLayout L;
    L.inst.resize(count);
    L.begin.assign((size_t)SP_COUNT * (count + 1), 0);
    const int g2 = sizeof(F) == 32 ? 0 : 1;
    // regime: a wide batch (msm_wide_batch, digits counted under the
    // single-MSM plans) fills the GPU by itself, so its instances take the
    // throughput plans (fewer buckets, longer runs)
    int regime = ctx->force_regime;
    if (regime < 0) {
        uint64_t e = 0;
        for (size_t i = 0; i < count; ++i) {
            const MsmPlan q = msm_plan_ctx(ctx, g2, jobs[i].n, (int)jobs[i].win_siz, MSM_REGIME_SINGLE);
            e += (uint64_t)jobs[i].n * (uint64_t)q.W;
        }
        regime = msm_wide_batch(g2, count, e) ? MSM_REGIME_BATCH : MSM_REGIME_SINGLE;
    }
    L.regime = regime;
    for (size_t i = 0; i < count; ++i) {
        MsmPlan p = msm_plan_ctx(ctx, g2, jobs[i].n, (int)jobs[i].win_siz, regime);
        MsmInst& d = L.inst[i];
        d.scalars = jobs[i].scalars;
        d.bases = jobs[i].bases;
        d.dest = jobs[i].dest;
        d.n = (uint32_t)jobs[i].n;
        d.c = (uint32_t)p.c;
        d.W = (uint32_t)p.W;
        d.NB = p.NB;
        d.nk = d.W * d.NB;
        d.K = p.K;
        d.g = p.g;
        d.G = p.NB / p.g;
        d.T = p.T;
        d.tile = p.tile;
        d.nd = d.n;
        d.Wd = d.W;
        d.Lg = 1;
        d.wb = 255 / d.W;   // balanced windows (c = ceil(255 / W) >= every width)
        d.wx = 255 % d.W;
        if (jobs[i].levels > 1) {
            // level table: every bucket set serves Lg digit windows over nd * Lg
            // virtual points; fewer bucket sets to reduce, larger buckets
            d.wb = d.c;   // the table's levels are 2^(c t) multiples: uniform windows
            d.wx = 0;
            d.Lg = jobs[i].levels;
            d.W = (d.Wd + d.Lg - 1) / d.Lg;
            d.n = d.nd * d.Lg;
            d.nk = d.W * d.NB;
            const uint64_t entries = (uint64_t)d.n * d.W;
            uint32_t K = 4;
            while (K < 4096 && entries / (2 * (uint64_t)K) >= ((uint64_t)1 << 18)) K <<= 1;
            d.K = K;
            uint32_t T = (512 + d.W - 1) / d.W;
            while (T > 1 && (d.n / T < 4096 || d.n / T < 4 * d.NB)) T >>= 1;
            d.T = T;
            d.tile = (d.n + T - 1) / T;
        }
        d.S = 1;   // window-sum partials: at most ~2 run sums per lane of 256
        while (d.S < 64 && d.G / d.S > 512) d.S *= 2;
        // single MSMs: one run per lane of 256, reduced from the buckets inside
        // the window-sum kernel (2g + 20 dependent additions per lane instead
        // of 2g in a run-sum pass and then up to 28), while the W * G / 256
        // workgroups of four waves fit one wave per SIMD (a second round of
        // waves would double the kernel's latency-bound time)
        d.fused = regime == MSM_REGIME_SINGLE && d.G % 256 == 0 && d.G / 256 <= 64 && (uint64_t)d.W * (d.G / 256) <= 256;
        if (d.fused) d.S = d.G / 256;
        uint64_t entries = (uint64_t)d.n * d.W;
        uint64_t sz[SP_COUNT];
        sz[SP_DIGIT] = d.nd;
        sz[SP_TILEWG] = (uint64_t)d.T * d.W;
        sz[SP_KEY] = (uint64_t)d.nk + 1;
        sz[SP_SCANTILE] = ((uint64_t)d.nk + MSM_SCAN_TILE) / MSM_SCAN_TILE;
        sz[SP_CHUNK] = (entries + d.K - 1) / d.K;
        sz[SP_BRED] = (uint64_t)d.W * d.G;
        sz[SP_WIN] = (uint64_t)d.W * d.S;
        sz[SP_ENTRY] = (entries + 63) & ~(uint64_t)63;   // 256-B aligned per instance (vector index loads)
        sz[SP_HIST] = (uint64_t)d.T * d.nk;
        d.tp = 0;   // two-pass flag and coarse groups: set below, once the batch's digit total is known
        sz[SP_CGROUP] = 0;
        for (int s = 0; s < SP_COUNT; ++s) L.begin[(size_t)s * (count + 1) + i + 1] = L.begin[(size_t)s * (count + 1) + i] + sz[s];
        if (d.NB > L.max_nb) L.max_nb = d.NB;
        uint64_t maxspan = (d.n + d.K - 1) / d.K + 1;   // a bucket holds at most n entries
        uint32_t lv = 0;
        if (maxspan > HOT)
            for (uint64_t stride = 1; stride < maxspan; stride *= K2) ++lv;
        if (lv > L.tree_levels) L.tree_levels = lv;
    }
    // two-pass scatter per instance, then the coarse-group space it needs
    uint64_t batch_entries = 0;
    for (const MsmInst& d : L.inst) batch_entries += (uint64_t)d.n * d.W;
    for (size_t i = 0; i < count; ++i) {
        MsmInst& d = L.inst[i];
        d.tp = msm_two_pass(d.n, d.NB, batch_entries) ? 1u : 0u;
        // coarse entries fit one word (index | fine key << 24 | sign << 31)
        // when point indices stay below 2^24 and a group has at most 128 buckets
        if (d.tp && d.n <= (1u << 24) && d.NB / MSM_COARSE <= 128) d.tp = 2;
        if (d.tp) L.two_pass = true;
        const size_t s = SP_CGROUP;
        L.begin[s * (count + 1) + i + 1] = L.begin[s * (count + 1) + i] + (d.tp ? (uint64_t)d.W * MSM_COARSE : 0);
    }
    for (int s = 0; s < SP_COUNT; ++s) L.tot[s] = L.begin[(size_t)s * (count + 1) + count];
    size_t off = 0;
    auto take = [&off](size_t bytes) { size_t o = (off + 255) & ~size_t(255); off = o + bytes; return o; };
    L.o_inst = take(count * sizeof(MsmInst));
    L.o_begin = take(L.begin.size() * 8);
    L.o_stats = take(count * 16);
    L.o_digits = take(L.tot[SP_ENTRY] * 4);
    L.o_hist = take(L.tot[SP_HIST] * 4);
    L.o_count = take(L.tot[SP_KEY] * 4);
    L.o_offs = take(L.tot[SP_KEY] * 4);
    L.o_tiles = take(2 * L.tot[SP_SCANTILE] * 8);
    L.o_sorted = take(L.tot[SP_ENTRY] * 4);
    L.o_coarse = take(L.two_pass ? L.tot[SP_ENTRY] * 8 : 0);
    const size_t raw = (size_t)raw_vec<F>() * 16;   // bytes per raw accumulator state
    L.o_chunk2key = take(L.tot[SP_CHUNK] * 4);
    L.o_bucket = take(L.tot[SP_KEY] * raw);
    L.o_carry = take(2 * L.tot[SP_CHUNK] * raw);             // head and tail partial per chunk
    L.o_segsum = take(2 * L.tot[SP_BRED] * raw);             // A_r, R_r per run
    L.o_win = take(2 * L.tot[SP_WIN] * sizeof(xyzz_t<F>));   // P_p, RS_p per part (R = 2^256 points, for the host)
    {
        std::map<const void*, uint64_t> need;   // source array -> points referenced (virtual points of level tables)
        for (size_t i = 0; i < count; ++i)
            if (jobs[i].form == 0) need[L.inst[i].bases] = std::max<uint64_t>(need[L.inst[i].bases], L.inst[i].n);
        std::map<const void*, uint32_t> task_of;
        uint64_t pts = 0;
        for (const auto& kv : need) {
            task_of[kv.first] = (uint32_t)L.conv.size();
            L.conv.push_back({kv.first, (void*)(uintptr_t)(pts * (g2 ? 128 : 64)), pts});
            pts += kv.second;
        }
        L.conv_total = pts;
        L.conv_of.assign(count, UINT32_MAX);   // UINT32_MAX: bases already converted
        for (size_t i = 0; i < count; ++i)
            if (jobs[i].form == 0) L.conv_of[i] = task_of[L.inst[i].bases];
        L.o_conv = take(pts * (g2 ? 128 : 64));
        L.o_convdesc = take(L.conv.size() * sizeof(MsmConvTask));
    }
    L.bytes = off + 256;
    return L;
}

// device stages 1-8 for a laid-out batch (all launches on ctx->stream; no host waits)
template <class F>
static void msm_device_stages(hzk_ctx* ctx, const MsmBatchLayout& L, char* base) {
    const size_t ninst = L.inst.size();
    MsmBatchDev b;
    b.inst = (const MsmInst*)(base + L.o_inst);
    b.begin = (const uint64_t*)(base + L.o_begin);
    b.ninst = (uint32_t)ninst;
    b.stats = (uint32_t*)(base + L.o_stats);
    MsmScalarBufs sb;
    sb.digits = (uint32_t*)(base + L.o_digits);
    sb.hist = (uint32_t*)(base + L.o_hist);
    sb.count = (uint32_t*)(base + L.o_count);
    sb.offs = (uint32_t*)(base + L.o_offs);
    sb.tiles = (uint64_t*)(base + L.o_tiles);
    sb.sorted = (uint32_t*)(base + L.o_sorted);
    sb.coarse = L.two_pass ? (uint2*)(base + L.o_coarse) : nullptr;
    uint32_t* chunk2key = (uint32_t*)(base + L.o_chunk2key);
    uint4* bucket = (uint4*)(base + L.o_bucket);
    uint4* carry = (uint4*)(base + L.o_carry);
    uint4* segsum = (uint4*)(base + L.o_segsum);
    xyzz_t<F>* win = (xyzz_t<F>*)(base + L.o_win);
    hipStream_t st = ctx->stream;

    if (L.conv_total) {
        TimedLaunch tl(ctx, "msm_convert");
        const uint64_t blocks = (L.conv_total + 255) / 256, cap = (uint64_t)ctx->num_cus * 32;
        const dim3 grid((unsigned)(blocks < cap ? blocks : cap));
        const MsmConvTask* tasks = (const MsmConvTask*)(base + L.o_convdesc);
        if constexpr (sizeof(F) == 32)
            hipLaunchKernelGGL(k_msm_convert_g1, grid, dim3(256), 0, st, tasks, (uint32_t)L.conv.size(), L.conv_total);
        else
            hipLaunchKernelGGL(k_msm_convert_g2, grid, dim3(256), 0, st, tasks, (uint32_t)L.conv.size(), L.conv_total);
    }
    msm_launch_scalar_stages(ctx, b, L.tot, L.max_nb, sb);
    {
        TimedLaunch tl(ctx, "msm_accumulate");
        uint64_t blocks = (L.tot[SP_CHUNK] + 255) / 256;
        uint64_t cap = (uint64_t)ctx->num_cus * 16;   // persistent grid: 16 waves per CU
        const dim3 grid((unsigned)(blocks < cap ? blocks : cap));
        if constexpr (sizeof(F) == 32)
            hipLaunchKernelGGL(k_msm_accumulate_g1, grid, dim3(256), 0, st, b, (const uint32_t*)sb.sorted, (const uint32_t*)sb.offs,
                               bucket, carry, chunk2key);
        else
            hipLaunchKernelGGL(k_msm_accumulate_g2, grid, dim3(256), 0, st, b, (const uint32_t*)sb.sorted, (const uint32_t*)sb.offs,
                               bucket, carry, chunk2key);
    }
    {
        TimedLaunch tl(ctx, "msm_fixup");
        uint32_t stride = 1;
        for (uint32_t lv = 0; lv < L.tree_levels; ++lv, stride *= K2)
            hipLaunchKernelGGL(k_msm_hot_tree<F>, dim3(std::min<unsigned>(blocks_for(L.tot[SP_CHUNK], 256), 1024)), dim3(256), 0, st, b,
                               (const uint32_t*)sb.offs, (const uint32_t*)chunk2key, stride, carry);
        hipLaunchKernelGGL(k_msm_fixup<F>, dim3(blocks_for(L.tot[SP_CHUNK], 256)), dim3(256), 0, st, b, (const uint32_t*)sb.offs,
                           (const uint32_t*)chunk2key, (const uint4*)carry, bucket);
    }
    {
        TimedLaunch tl(ctx, "msm_bucket_reduce");
        bool all_fused = true;
        for (const MsmInst& d : L.inst) all_fused = all_fused && d.fused;
        if (!all_fused)
            hipLaunchKernelGGL(k_bucket_segments<F>, dim3(blocks_for(L.tot[SP_BRED], 128)), dim3(128), 0, st, b, (const uint4*)bucket,
                               (const uint32_t*)sb.offs, segsum);
        constexpr size_t WW = acc29_t<F>::RAW_WORDS * 4;   // LDS bytes per thread slot
        if (L.regime == MSM_REGIME_BATCH)
            hipLaunchKernelGGL((k_window_sum<F, 64>), dim3((unsigned)L.tot[SP_WIN]), dim3(64), WW * 64, st, b, (const uint4*)segsum,
                               (const uint4*)bucket, (const uint32_t*)sb.offs, win);
        else
            hipLaunchKernelGGL((k_window_sum<F, 256>), dim3((unsigned)L.tot[SP_WIN]), dim3(256), WW * 256, st, b, (const uint4*)segsum,
                               (const uint4*)bucket, (const uint32_t*)sb.offs, win);
    }
}

template <class F>
static void msm_launch_combine(hzk_ctx* ctx, const MsmBatchLayout& L, char* base) {
    MsmBatchDev b;
    b.inst = (const MsmInst*)(base + L.o_inst);
    b.begin = (const uint64_t*)(base + L.o_begin);
    b.ninst = (uint32_t)L.inst.size();
    b.stats = (uint32_t*)(base + L.o_stats);
    uint32_t maxw = 1;
    for (const MsmInst& d : L.inst) maxw = std::max(maxw, d.W);
    uint4* tsum = (uint4*)(base + L.o_segsum);   // W raw states per instance fit its run-sum range
    TimedLaunch tl(ctx, "msm_combine");
    hipLaunchKernelGGL(k_msm_combine_windows<F>, dim3((unsigned)(L.inst.size() * maxw)), dim3(64), 0, ctx->stream, b,
                       (const xyzz_t<F>*)(base + L.o_win), maxw, tsum);
    hipLaunchKernelGGL(k_msm_combine_horner<F>, dim3(blocks_for(L.inst.size(), 64)), dim3(64), 0, ctx->stream, b,
                       (const uint4*)tsum);
}

template <class F>
static std::string shape_key(const MsmBatchLayout& L, bool combine) {
    std::string k(sizeof(F) == 32 ? "g1:" : "g2:");
    k += L.regime == MSM_REGIME_BATCH ? "b:" : "s:";
    if (combine) k += "dev:";
    k += std::to_string(L.conv.size()) + "/" + std::to_string(L.conv_total) + ":";
    for (const MsmInst& d : L.inst) {
        char buf[64];
        snprintf(buf, sizeof buf, "%u/%u/%u/%u/%u/%u;", d.n, d.c, d.K, d.g, d.T, d.Lg);
        k += buf;
    }
    return k;
}

// ext_scratch != null: run on that device buffer (ext_bytes long) instead of
// the context's arena, with pageable host staging; nothing is allocated on
// the device or pinned on the host (no graph capture in this mode)
template <class F, class HostPt>
int msm_batch_run(hzk_ctx* ctx, const MsmJob* jobs, size_t count, int dests_on_host, int capture, void* ext_scratch = nullptr,
                  size_t ext_bytes = 0, int device_combine = 0) {
    HZK_CHECK(ctx, count > 0 && jobs, HZK_ERR_INVALID, "msm: empty batch");
    for (size_t i = 0; i < count; ++i) {
        HZK_CHECK(ctx, jobs[i].n > 0 && jobs[i].n < (1ull << 31), HZK_ERR_INVALID, "msm: n=%zu out of range [1, 2^31)", jobs[i].n);
        HZK_CHECK(ctx, jobs[i].levels <= 1 || (jobs[i].win_siz && jobs[i].levels <= (255 + jobs[i].win_siz - 1) / jobs[i].win_siz),
                  HZK_ERR_INVALID, "msm: a level table needs win_siz and at most ceil(255 / win_siz) levels (%u)", jobs[i].levels);
        HZK_CHECK(ctx, (uint64_t)jobs[i].n * (jobs[i].levels ? jobs[i].levels : 1) < (1ull << 31), HZK_ERR_UNSUPPORTED,
                  "msm: n=%zu x %u levels exceeds 2^31 table points", jobs[i].n, jobs[i].levels);
        HZK_CHECK(ctx, jobs[i].scalars && jobs[i].bases && jobs[i].dest, HZK_ERR_INVALID, "msm: null pointer in job %zu", i);
        HZK_CHECK(ctx, jobs[i].form <= 1, HZK_ERR_INVALID, "msm: base form %u", jobs[i].form);
        HZK_CHECK(ctx, jobs[i].win_siz == 0 || (jobs[i].win_siz >= (uint32_t)MSM_MIN_C && jobs[i].win_siz <= (uint32_t)MSM_MAX_C),
                  HZK_ERR_INVALID, "msm: win_siz=%u unsupported (0 or %d..%d)", jobs[i].win_siz, MSM_MIN_C, MSM_MAX_C);
    }
    HZK_CHECK(ctx, !(device_combine && dests_on_host), HZK_ERR_INVALID, "msm: the device-side combine needs device dests");
    msm_prepare();
    MsmBatchLayout L = msm_layout<F>(ctx, jobs, count);
    for (size_t i = 0; i < count; ++i)   // bucket offsets are 32-bit
        HZK_CHECK(ctx, (uint64_t)L.inst[i].n * L.inst[i].W <= 0xffff0000ull, HZK_ERR_UNSUPPORTED,
                  "msm: n=%u with %u windows exceeds 2^32 bucket entries", L.inst[i].n, L.inst[i].W);
    HZK_CHECK(ctx, (size_t)acc29_t<F>::RAW_WORDS * 4 * 256 <= ctx->lds_per_block, HZK_ERR_UNSUPPORTED,
              "msm: the window sums need %zu B of LDS per workgroup, device allows %zu", (size_t)acc29_t<F>::RAW_WORDS * 4 * 256,
              ctx->lds_per_block);
    HZK_CHECK(ctx, (size_t)L.max_nb * 4 <= ctx->lds_per_block, HZK_ERR_UNSUPPORTED,
              "msm: a %u-bucket window needs %u B of LDS per workgroup, device allows %zu", L.max_nb, L.max_nb * 4,
              ctx->lds_per_block);
    int rc = 0;
    if (ext_scratch) {
        HZK_CHECK(ctx, L.bytes <= ext_bytes && !capture, HZK_ERR_UNSUPPORTED,
                  "msm: this MSM needs %zu bytes of scratch, the static buffer holds %zu", L.bytes, ext_bytes);
    } else {
        rc = ensure_scratch(ctx, L.bytes);
        if (rc) return rc;
    }
    const size_t conv_bytes = L.conv.size() * sizeof(MsmConvTask);
    const size_t host_bytes = count * sizeof(MsmInst) + L.begin.size() * 8 + 2 * L.tot[SP_WIN] * sizeof(xyzz_t<F>) + count * 16 +
                              conv_bytes + 1024;
    char* pin;
    int slot = -1;
    std::vector<char> pageable;
    if (ext_scratch) {
        pageable.resize(host_bytes);
        pin = pageable.data();
    } else {
        rc = stage_acquire(ctx, host_bytes, &pin, &slot);
        if (rc) return rc;
    }
    char* base = ext_scratch ? (char*)ext_scratch : (char*)ctx->scratch;
    hipStream_t st = ctx->stream;
    // G1: instances read their converted base arrays in the scratch
    for (MsmConvTask& t : L.conv) t.dst = base + L.o_conv + (size_t)(uintptr_t)t.dst;
    for (size_t i = 0; i < L.conv_of.size(); ++i)
        if (L.conv_of[i] != UINT32_MAX) L.inst[i].bases = L.conv[L.conv_of[i]].dst;
    // descriptors and work-space starts
    std::memcpy(pin, L.inst.data(), count * sizeof(MsmInst));
    std::memcpy(pin + count * sizeof(MsmInst), L.begin.data(), L.begin.size() * 8);
    HZK_HIP(ctx, hipMemcpyAsync(base + L.o_inst, pin, count * sizeof(MsmInst), hipMemcpyHostToDevice, st));
    HZK_HIP(ctx, hipMemcpyAsync(base + L.o_begin, pin + count * sizeof(MsmInst), L.begin.size() * 8, hipMemcpyHostToDevice, st));
    char* pconv = pin + count * sizeof(MsmInst) + L.begin.size() * 8 + 2 * L.tot[SP_WIN] * sizeof(xyzz_t<F>) + count * 16;
    if (conv_bytes) {
        std::memcpy(pconv, L.conv.data(), conv_bytes);
        HZK_HIP(ctx, hipMemcpyAsync(base + L.o_convdesc, pconv, conv_bytes, hipMemcpyHostToDevice, st));
    }

    if (capture) {
        std::string key = shape_key<F>(L, device_combine != 0);
        hzk_graph_entry& g = ctx->graphs[key];
        if (!g.exec || g.generation != ctx->scratch_generation) {
            if (g.exec) { (void)hipGraphExecDestroy(g.exec); g.exec = nullptr; }
            hipGraph_t graph;
            bool timing = ctx->timing;
            ctx->timing = false;
            HZK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            msm_device_stages<F>(ctx, L, base);
            if (device_combine) msm_launch_combine<F>(ctx, L, base);
            hipError_t e = hipStreamEndCapture(st, &graph);
            ctx->timing = timing;
            HZK_HIP(ctx, e);
            e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HZK_HIP(ctx, e);
            g.generation = ctx->scratch_generation;
        }
        TimedLaunch tl(ctx, "msm_graph");
        HZK_HIP(ctx, hipGraphLaunch(g.exec, st));
    } else {
        msm_device_stages<F>(ctx, L, base);
        if (device_combine) msm_launch_combine<F>(ctx, L, base);
    }
    HZK_HIP(ctx, hipGetLastError());
    if (device_combine) {   // results are written by the device; nothing comes back to the host
        if (slot >= 0) return stage_fence(ctx, slot);
        HZK_HIP(ctx, hipStreamSynchronize(st));   // pageable staging (no-alloc scratch): wait for the copies
        return 0;
    }
    char* hwin = pin + count * sizeof(MsmInst) + L.begin.size() * 8;
    uint32_t* hstats = (uint32_t*)(hwin + 2 * L.tot[SP_WIN] * sizeof(xyzz_t<F>));
    HZK_HIP(ctx, hipMemcpyAsync(hwin, base + L.o_win, 2 * L.tot[SP_WIN] * sizeof(xyzz_t<F>), hipMemcpyDeviceToHost, st));
    HZK_HIP(ctx, hipMemcpyAsync(hstats, base + L.o_stats, count * 16, hipMemcpyDeviceToHost, st));
    if (slot >= 0) {
        rc = stage_fence(ctx, slot);
        if (rc) return rc;
    }
    HZK_HIP(ctx, hipStreamSynchronize(st));

    // window values on the host, all (instance, window) pairs in parallel:
    // T_j = sum_p P_p + (g Gs) * sum_p p RS_p, the last sum as the sum of the
    // suffix sums of RS over p >= 1
    const uint64_t* wbeg = L.begin.data() + (size_t)SP_WIN * (count + 1);
    std::vector<uint64_t> tbeg(count + 1, 0);
    for (size_t i = 0; i < count; ++i) tbeg[i + 1] = tbeg[i] + L.inst[i].W;
    std::vector<HostPt> T(tbeg[count]);
    parallel_for((long)tbeg[count], [&](long q) {
        const size_t i = std::upper_bound(tbeg.begin(), tbeg.end(), (uint64_t)q) - tbeg.begin() - 1;
        const MsmInst& d = L.inst[i];
        const size_t j = (size_t)q - tbeg[i];
        auto pt = [&](uint32_t p, int which) {
            return HostPt::load_xyzz(hwin + (2 * (wbeg[i] + j * d.S + p) + which) * sizeof(xyzz_t<F>));
        };
        HostPt acc = HostPt::inf(), suf = HostPt::inf(), y = HostPt::inf();
        for (uint32_t p = 0; p < d.S; ++p) acc = acc.add(pt(p, 0));
        for (uint32_t p = d.S - 1; p >= 1; --p) {
            suf = suf.add(pt(p, 1));
            y = y.add(suf);
        }
        if (d.S > 1) {
            for (uint32_t k = 1; k < d.G / d.S * d.g; k <<= 1) y = y.dbl();
            acc = acc.add(y);
        }
        T[q] = acc;
    });
    // Horner over windows, instances in parallel
    std::vector<char> out(count * 3 * sizeof(F));
    parallel_for((long)count, [&](long i) {
        const MsmInst& d = L.inst[i];
        HostPt acc = HostPt::inf();
        for (int j = (int)d.W - 1; j >= 0; --j) {
            // a bucket set spans Lg digit windows of c bits; plain windows are wb or wb + 1 bits
            const uint32_t shift = d.Lg > 1 ? d.c * d.Lg : d.wb + ((uint32_t)j < d.wx ? 1u : 0u);
            for (uint32_t q = 0; q < shift; ++q) acc = acc.dbl();
            acc = acc.add(T[tbeg[i] + j]);
        }
        acc.store_jacobian(&out[i * 3 * sizeof(F)]);
    });
    for (size_t i = 0; i < count; ++i) {
        if (dests_on_host) std::memcpy(jobs[i].dest, &out[i * 3 * sizeof(F)], 3 * sizeof(F));
        else HZK_HIP(ctx, hipMemcpyAsync(jobs[i].dest, &out[i * 3 * sizeof(F)], 3 * sizeof(F), hipMemcpyHostToDevice, st));
    }
    if (!dests_on_host) HZK_HIP(ctx, hipStreamSynchronize(st));
    const MsmInst& d0 = L.inst[count - 1];
    ctx->last_msm[0] = hstats[4 * (count - 1)];
    ctx->last_msm[1] = hstats[4 * (count - 1) + 1];
    ctx->last_msm[2] = hstats[4 * (count - 1) + 2];
    ctx->last_msm[3] = d0.c;
    ctx->last_msm[4] = d0.W;
    return 0;
}

template <class F, class HostPt>
int msm_run(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest, uint32_t levels) {
    HZK_CHECK(ctx, win_siz <= (size_t)MSM_MAX_C, HZK_ERR_INVALID, "msm: win_siz=%zu unsupported (0 or %d..%d)", win_siz,
              MSM_MIN_C, MSM_MAX_C);
    MsmJob j{scalars, bases, n, dest, (uint32_t)win_siz, levels};
    return msm_batch_run<F, HostPt>(ctx, &j, 1, 0, 0);
}

}  // namespace hzk
