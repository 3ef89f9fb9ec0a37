// Group-generic MSM kernels and driver (G1 over Fq, G2 over Fq2); included
// by msm_g1.hip and msm_g2.hip, one instantiation per translation unit. The
// stage list is in msm.hip.
#pragma once
#include <cstring>
#include "common.hpp"
#include "curve.hpp"
#include "host_ec.hpp"
#include "msm.hpp"

namespace hzk {

static constexpr int K2 = 16;   // merge fan-in

// ------------------------------------------------------------------ accumulate
// index of the key whose [segbase[k], segbase[k+1]) contains s
HZK_DEV uint32_t find_key(const uint32_t* __restrict__ segbase, uint32_t nk, uint32_t s) {
    uint32_t lo = 0, hi = nk;   // invariant: segbase[lo] <= s < segbase[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (segbase[mid] <= s) lo = mid; else hi = mid;
    }
    return lo;
}

template <class F>
__global__ void __launch_bounds__(256) k_msm_accumulate(const affine_t<F>* __restrict__ bases, const uint32_t* __restrict__ sorted,
                                                        const uint32_t* __restrict__ offs, const uint32_t* __restrict__ segbase,
                                                        uint32_t nk, uint32_t nseg, uint32_t K,
                                                        xyzz_t<F>* __restrict__ part, uint32_t* __restrict__ seg2key) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint32_t k = find_key(segbase, nk, s);
    uint32_t e = offs[k] + (s - segbase[k]) * K;
    uint32_t e1 = min(e + K, offs[k + 1]);
    xyzz_t<F> acc = xyzz_t<F>::inf();
    uint32_t idx = sorted[e];
    affine_t<F> P = bases[idx & 0x7fffffffu];
    for (; e < e1; ++e) {
        bool neg = idx >> 31;
        affine_t<F> Q = P;
        if (e + 1 < e1) {   // prefetch the next base while this one is added
            idx = sorted[e + 1];
            P = bases[idx & 0x7fffffffu];
        }
        if (neg) Q.y = Q.y.neg();
        acc.add_affine(Q);
    }
    part[s] = acc;
    seg2key[s] = k;
}

// keys with 2..K2 segments: one lane per key sums them into the first slot
template <class F>
__global__ void __launch_bounds__(256) k_msm_merge_small(const uint32_t* __restrict__ segbase, uint32_t nk,
                                                         xyzz_t<F>* __restrict__ part) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    uint32_t b0 = segbase[k], cnt = segbase[k + 1] - b0;
    if (cnt < 2 || cnt > K2) return;
    xyzz_t<F> acc = part[b0];
    for (uint32_t i = 1; i < cnt; ++i) acc.add(part[b0 + i]);
    part[b0] = acc;
}

// keys with more than K2 segments (oversized buckets): one level of an
// in-place K2-ary tree over the key's segment sums
template <class F>
__global__ void __launch_bounds__(256) k_msm_merge(const uint32_t* __restrict__ segbase, const uint32_t* __restrict__ seg2key,
                                                   uint32_t nseg, uint32_t stride, xyzz_t<F>* __restrict__ part) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint32_t k = seg2key[s];
    uint32_t b0 = segbase[k], cnt = segbase[k + 1] - b0, rel = s - b0;
    if (cnt <= K2 || cnt <= stride || rel % (stride * K2) != 0) return;
    xyzz_t<F> acc = part[s];
    for (int i = 1; i < K2; ++i) {
        uint32_t r = rel + i * stride;
        if (r >= cnt) break;
        acc.add(part[b0 + r]);
    }
    part[s] = acc;
}

// ------------------------------------------------------------------ bucket reduction
template <class F>
HZK_DEV xyzz_t<F> bucket_sum(const xyzz_t<F>* __restrict__ part, const uint32_t* __restrict__ offs,
                              const uint32_t* __restrict__ segbase, uint32_t key) {
    if (offs[key + 1] == offs[key]) return xyzz_t<F>::inf();
    return part[segbase[key]];
}

// lane per (window, run of g buckets): T = sum_{d in run} d * B_d
template <class F>
__global__ void __launch_bounds__(128) k_bucket_segments(const xyzz_t<F>* __restrict__ part, const uint32_t* __restrict__ offs,
                                                         const uint32_t* __restrict__ segbase, uint32_t W, uint32_t NB, uint32_t g,
                                                         xyzz_t<F>* __restrict__ segsum) {
    uint32_t G = NB / g;
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= W * G) return;
    uint32_t j = t / G, sg = t % G;
    uint32_t kb = j * NB + sg * g;
    xyzz_t<F> run = xyzz_t<F>::inf(), acc = xyzz_t<F>::inf();
    for (int b = (int)g - 1; b >= 0; --b) {
        run.add(bucket_sum<F>(part, offs, segbase, kb + b));
        acc.add(run);
    }
    // acc = sum (d - sg*g) B_d ; add (sg*g) * run
    uint32_t w = sg * g;
    if (w) {
        xyzz_t<F> m = xyzz_t<F>::inf();
        for (int bit = 31 - __clz(w); bit >= 0; --bit) {
            m = m.dbl();
            if ((w >> bit) & 1) m.add(run);
        }
        acc.add(m);
    }
    segsum[t] = acc;
}

// block per window: tree sum of the window's segment sums
template <class F>
__global__ void __launch_bounds__(256) k_window_sum(const xyzz_t<F>* __restrict__ segsum, uint32_t G, xyzz_t<F>* __restrict__ winsum) {
    __shared__ xyzz_t<F> sh[256];
    uint32_t j = blockIdx.x;
    xyzz_t<F> acc = xyzz_t<F>::inf();
    for (uint32_t s = threadIdx.x; s < G; s += 256) acc.add(segsum[j * G + s]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            xyzz_t<F> a = sh[threadIdx.x];
            a.add(sh[threadIdx.x + off]);
            sh[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) winsum[j] = sh[0];
}

// ------------------------------------------------------------------ driver
template <class F, class HostPt>
int msm_run(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest) {
    HZK_CHECK(ctx, n > 0 && n < (1ull << 31), HZK_ERR_INVALID, "msm: n=%zu out of range [1, 2^31)", n);
    HZK_CHECK(ctx, scalars && bases && dest, HZK_ERR_INVALID, "msm: null pointer");
    HZK_CHECK(ctx, win_siz == 0 || (win_siz >= MSM_MIN_C && win_siz <= MSM_MAX_C), HZK_ERR_INVALID,
              "msm: win_siz=%zu unsupported (0 or %d..%d)", win_siz, MSM_MIN_C, MSM_MAX_C);
    MsmPlan p = msm_plan(n, win_siz ? (int)win_siz : ctx->force_window, ctx->force_seg, ctx->lds_per_block);
    HZK_CHECK(ctx, msm_lds_bytes(p) <= ctx->lds_per_block, HZK_ERR_UNSUPPORTED,
              "msm: window %d needs %zu B of LDS per workgroup, device allows %zu", p.c, msm_lds_bytes(p), ctx->lds_per_block);
    HZK_CHECK(ctx, p.g >= 1 && (p.g & (p.g - 1)) == 0 && p.g <= p.NB, HZK_ERR_INVALID, "msm: bad segment size %u", p.g);
    const uint32_t nk = (uint32_t)p.W * p.NB;
    const size_t Lmax = n * (size_t)p.W;
    const size_t Smax = Lmax / p.K + nk;
    const uint32_t ntiles = msm_scan_tiles(nk);
    const uint32_t G = p.NB / p.g;

    size_t need = 0;
    auto add = [&need](size_t b) { need = ((need + 255) & ~size_t(255)) + b; };
    add(Lmax * 4ull);                           // digits (window-major)
    add((size_t)p.T * nk * 4ull);               // per-tile histograms
    add(nk * 4ull);                             // count
    add((nk + 1) * 4ull);                       // offs
    add((nk + 1) * 4ull);                       // segbase
    add(2ull * ntiles * 8ull);                  // scan tile sums + maxima
    add(64);                                    // stats
    add(Lmax * 4ull);                           // sorted
    add(Smax * 4ull);                           // seg2key
    add(Smax * sizeof(xyzz_t<F>));              // part
    add((size_t)p.W * G * sizeof(xyzz_t<F>));   // segsum
    add((size_t)p.W * sizeof(xyzz_t<F>));       // winsum
    int rc = ensure_scratch(ctx, need + 4096);
    if (rc) return rc;
    rc = ensure_pinned(ctx, 64 + p.W * sizeof(xyzz_t<F>));
    if (rc) return rc;
    Arena ar{(char*)ctx->scratch, ctx->scratch_bytes};
    uint32_t* digits = (uint32_t*)ar.take(Lmax * 4ull);
    uint32_t* hist = (uint32_t*)ar.take((size_t)p.T * nk * 4ull);
    uint32_t* count = (uint32_t*)ar.take(nk * 4ull);
    uint32_t* offs = (uint32_t*)ar.take((nk + 1) * 4ull);
    uint32_t* segbase = (uint32_t*)ar.take((nk + 1) * 4ull);
    uint64_t* tiles = (uint64_t*)ar.take(2ull * ntiles * 8ull);
    uint32_t* stats = (uint32_t*)ar.take(64);
    uint32_t* sorted = (uint32_t*)ar.take(Lmax * 4ull);
    uint32_t* seg2key = (uint32_t*)ar.take(Smax * 4ull);
    xyzz_t<F>* part = (xyzz_t<F>*)ar.take(Smax * sizeof(xyzz_t<F>));
    xyzz_t<F>* segsum = (xyzz_t<F>*)ar.take((size_t)p.W * G * sizeof(xyzz_t<F>));
    xyzz_t<F>* winsum = (xyzz_t<F>*)ar.take((size_t)p.W * sizeof(xyzz_t<F>));

    hipStream_t st = ctx->stream;
    msm_launch_sort(ctx, p, scalars, (uint32_t)n, digits, hist, count, tiles, ntiles, stats, offs, segbase);
    HZK_HIP(ctx, hipGetLastError());
    uint32_t* hstats = (uint32_t*)ctx->pinned;
    HZK_HIP(ctx, hipMemcpyAsync(hstats, stats, 16, hipMemcpyDeviceToHost, st));
    msm_launch_scatter(ctx, p, digits, (uint32_t)n, hist, offs, sorted);
    HZK_HIP(ctx, hipGetLastError());
    HZK_HIP(ctx, hipStreamSynchronize(st));
    const uint32_t L = hstats[0], S = hstats[1], maxc = hstats[2];
    HZK_CHECK(ctx, L <= Lmax && S <= Smax, HZK_ERR_INVALID, "msm: internal size check failed (L=%u S=%u)", L, S);
    ctx->last_msm[0] = L; ctx->last_msm[1] = S; ctx->last_msm[2] = maxc;
    ctx->last_msm[3] = (uint32_t)p.c; ctx->last_msm[4] = (uint32_t)p.W;

    if (S > 0) {
        {
            TimedLaunch tl(ctx, "msm_accumulate");
            hipLaunchKernelGGL(k_msm_accumulate<F>, dim3(blocks_for(S, 256)), dim3(256), 0, st, (const affine_t<F>*)bases,
                               sorted, offs, segbase, nk, S, p.K, part, seg2key);
        }
        uint32_t maxseg = (maxc + p.K - 1) / p.K;
        if (maxseg > 1) {
            TimedLaunch tl(ctx, "msm_merge");
            hipLaunchKernelGGL(k_msm_merge_small<F>, dim3(blocks_for(nk, 256)), dim3(256), 0, st, segbase, nk, part);
        }
        for (uint32_t stride = 1; maxseg > (uint32_t)K2 && stride < maxseg; stride *= K2) {
            TimedLaunch tl(ctx, "msm_merge");
            hipLaunchKernelGGL(k_msm_merge<F>, dim3(blocks_for(S, 256)), dim3(256), 0, st, segbase, seg2key, S, stride, part);
        }
    }
    {
        TimedLaunch tl(ctx, "msm_bucket_reduce");
        hipLaunchKernelGGL(k_bucket_segments<F>, dim3(blocks_for((size_t)p.W * G, 128)), dim3(128), 0, st,
                           (const xyzz_t<F>*)part, offs, segbase, (uint32_t)p.W, p.NB, p.g, segsum);
        hipLaunchKernelGGL(k_window_sum<F>, dim3(p.W), dim3(256), 0, st, (const xyzz_t<F>*)segsum, G, winsum);
    }
    HZK_HIP(ctx, hipGetLastError());
    char* hwin = (char*)ctx->pinned + 64;
    HZK_HIP(ctx, hipMemcpyAsync(hwin, winsum, p.W * sizeof(xyzz_t<F>), hipMemcpyDeviceToHost, st));
    HZK_HIP(ctx, hipStreamSynchronize(st));

    // Horner over windows on the host
    HostPt acc = HostPt::inf();
    for (int j = p.W - 1; j >= 0; --j) {
        for (int b = 0; b < p.c; ++b) acc = acc.dbl();
        acc = acc.add(HostPt::load_xyzz(hwin + j * sizeof(xyzz_t<F>)));
    }
    char out[192];
    acc.store_jacobian(out);
    HZK_HIP(ctx, hipMemcpyAsync(dest, out, 3 * sizeof(F), hipMemcpyHostToDevice, st));
    HZK_HIP(ctx, hipStreamSynchronize(st));
    return 0;
}

}  // namespace hzk
