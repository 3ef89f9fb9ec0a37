// Fr number-theoretic transforms on gfx950: NTT, inverse NTT and coset NTT.
//
// Replaces number_theory_transformation (include/operator/poly.cuh:65-73:
// one launch and one full HBM read+write per radix-2 stage, half the lanes
// idle, then a bit-reverse launch and a scaling launch) while keeping its
// semantics exactly: natural-order input and output,
//   out[k] = factor * sum_j in[j] * w^(j k),   w = omega_powers[1],
// with omega_powers[e] = w^e for e < deg/2 (the table layout of ntt_setup,
// include/protocol/groth16-operator.cuh:6-32) and the same decimation-in-
// frequency butterfly (a, b) -> (a + b, (a - b) w^((j mod h) 2^s)).
//
// Here the log2(deg) stages are grouped into passes: each workgroup stages a
// 2^11-element tile (64 KiB) in LDS — 2^R butterfly rows by G = 2^(11-R)
// columns that are contiguous in HBM, so loads and stores stay coalesced —
// runs R stages there, and writes the tile back. 2^20 takes two passes
// (R = 9, then the last 11 stages, whose rows are contiguous). One more pass
// applies the bit-reverse permutation fused with the scale factor.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>
#include "common.hpp"
#include "field.hpp"
#include "host_ec.hpp"

namespace hzk {

static constexpr int NTT_TILE_BITS = 11;
// 1024 threads (16 waves) per 64 KiB tile: the LDS allows two tiles per CU,
// so the workgroup size sets the waves available to hide LDS and twiddle-load
// latency (256 threads: 41.6 G butterflies/s batched, 12.5 one transform at a
// time; 1024: 48.5 and 18.6)
static constexpr int NTT_THREADS = 1024;
static constexpr int NTT_MAX_LOG = 26;

// one tile of a pass of stages [s0, s0 + R) over tiles of 2^R rows x 2^gbits columns
HZK_DEV void ntt_tile(fr_t* __restrict__ x, const fr_t* __restrict__ tw, int k, int s0, int R, int gbits, uint32_t b, fr_t* sh) {
    const int low = k - s0 - R;
    const uint32_t G = 1u << gbits;
    const uint32_t T = 1u << (R + gbits);
    const uint32_t lchunks = (1u << low) >> gbits;
    const uint32_t hi = b / lchunks, lc = b % lchunks;
    const size_t base = ((size_t)hi << (k - s0)) | ((size_t)lc << gbits);
    for (uint32_t e = threadIdx.x; e < T; e += NTT_THREADS) {
        uint32_t m = e >> gbits, g = e & (G - 1);
        sh[e] = x[base + ((size_t)m << low) + g];
    }
    __syncthreads();
    for (int st = 0; st < R; ++st) {
        const int s = s0 + st;
        const int mp = R - 1 - st;
        const uint32_t mmask = (1u << mp) - 1;
        for (uint32_t bf = threadIdx.x; bf < T / 2; bf += NTT_THREADS) {
            uint32_t g = bf & (G - 1), q = bf >> gbits;
            uint32_t m0 = ((q >> mp) << (mp + 1)) | (q & mmask);
            uint32_t i0 = (m0 << gbits) | g, i1 = ((m0 | (1u << mp)) << gbits) | g;
            size_t e = ((((size_t)(m0 & mmask)) << low) | ((size_t)lc << gbits | g)) << s;
            fr_t a = sh[i0], c = sh[i1];
            sh[i0] = a + c;
            sh[i1] = (a - c) * tw[e];
        }
        __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < T; e += NTT_THREADS) {
        uint32_t m = e >> gbits, g = e & (G - 1);
        x[base + ((size_t)m << low) + g] = sh[e];
    }
}

__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass(fr_t* __restrict__ x, const fr_t* __restrict__ tw, int k, int s0,
                                                          int R, int gbits) {
    extern __shared__ fr_t sh[];
    ntt_tile(x, tw, k, s0, R, gbits, blockIdx.x, sh);
}

// ------------------------------------------------------------------ batched transforms
// one (transform, pass) of a round: its tiles occupy [tile_begin, next tile_begin)
struct NttPassDesc {
    fr_t* x;
    const fr_t* tw;
    int k, s0, R, gbits;
    uint64_t tile_begin;
};
// one transform for the element-wise stages: its elements occupy [begin, next begin)
struct NttElemDesc {
    fr_t* x;
    const fr_t* factor;   // bit-reverse scale (1/n for inverses) or null
    const fr_t* pre;      // coset powers applied before the passes, or null
    const fr_t* post;     // coset powers applied after the permutation, or null
    int k;
    uint64_t begin;
};

template <class D>
HZK_DEV uint32_t find_desc(const D* d, uint32_t nd, uint64_t gid) {
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (d[mid].tile_begin <= gid) lo = mid; else hi = mid;
    }
    return lo;
}
HZK_DEV uint32_t find_elem(const NttElemDesc* d, uint32_t nd, uint64_t gid) {
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (d[mid].begin <= gid) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(NTT_THREADS) k_ntt_pass_batch(const NttPassDesc* __restrict__ pd, uint32_t npd) {
    extern __shared__ fr_t sh[];
    uint32_t i = find_desc(pd, npd, blockIdx.x);
    const NttPassDesc& d = pd[i];
    ntt_tile(d.x, d.tw, d.k, d.s0, d.R, d.gbits, (uint32_t)(blockIdx.x - d.tile_begin), sh);
}

__global__ void __launch_bounds__(256) k_ntt_pre_batch(const NttElemDesc* __restrict__ ed, uint32_t ne, uint64_t total) {
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const NttElemDesc& d = ed[find_elem(ed, ne, gid)];
    if (!d.pre) return;
    uint64_t j = gid - d.begin;
    d.x[j] = d.x[j] * d.pre[j];
}

__global__ void __launch_bounds__(256) k_bitrev_batch(const NttElemDesc* __restrict__ ed, uint32_t ne, uint64_t total) {
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const NttElemDesc& d = ed[find_elem(ed, ne, gid)];
    size_t i = gid - d.begin;
    size_t r = d.k ? (size_t)(__brev((uint32_t)i) >> (32 - d.k)) : 0;
    if (i > r) return;
    fr_t a = d.x[i], c = d.x[r];
    if (d.factor) {
        fr_t f = *d.factor;
        a = a * f;
        c = c * f;
    }
    if (d.post) {
        c = c * d.post[i];
        a = a * d.post[r];
    }
    d.x[i] = c;
    d.x[r] = a;
}

// bit-reverse permutation (k bits) fused with an optional scale factor
__global__ void __launch_bounds__(256) k_bitrev_scale(fr_t* __restrict__ x, size_t n, int k, const fr_t* __restrict__ factor) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t r = (size_t)(__brev((uint32_t)i) >> (32 - k));
    if (i > r) return;
    fr_t a = x[i], c = x[r];
    if (factor) {
        fr_t f = *factor;
        a = a * f;
        c = c * f;
    }
    x[i] = c;
    x[r] = a;
}

// x[j] *= tab[j] (* scale)
__global__ void __launch_bounds__(256) k_mul_table(fr_t* __restrict__ x, const fr_t* __restrict__ tab, size_t n,
                                                   const fr_t* __restrict__ scale) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t v = x[i] * tab[i];
    if (scale) v = v * *scale;
    x[i] = v;
}

// out[e] = w^e for e < n, from sq[b] = w^(2^b)
__global__ void __launch_bounds__(256) k_powers(fr_t* __restrict__ out, size_t n, const fr_t* __restrict__ sq) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    fr_t acc = fr_t::one();
    for (int b = 0; (e >> b) != 0; ++b)
        if ((e >> b) & 1) acc = acc * sq[b];
    out[e] = acc;
}

// ------------------------------------------------------------------ host side
// Fr constants in Montgomery form: primitive 2^k-th roots of unity
// w_k = 5^((r-1)/2^k) (5 generates Fr*, alt_bn128_init.cu:50) and inverses
// are derived on the host from w_28, the value the reference stores as
// root_of_unity (alt_bn128_init.cu:51).
static host::Fr fr_from_u64(uint64_t v) {
    host::Fr t{{v, 0, 0, 0}};
    host::Fr r2;
    std::memcpy(r2.v, host::FR.r2, 32);
    return t * r2;
}
static host::Fr root_of_unity(int k) {
    // w_28 = 5^t, t = (r-1)/2^28
    static const uint64_t t[4] = {0x9b9709143e1f593fULL, 0x181585d2833e8487ULL, 0x131a029b85045b68ULL, 0x000000030644e72eULL};
    host::Fr w = fr_from_u64(5).pow(t);
    for (int i = 28; i > k; --i) w = w.sqr();
    return w;
}

// per transform size: forward / inverse twiddles (n/2 each), coset powers
// g^j and g^-j (n each) and 1/n; built on first use, kept for the context
struct NttTables {
    fr_t* tw[2] = {nullptr, nullptr};
    fr_t* coset[2] = {nullptr, nullptr};
    fr_t* ninv = nullptr;
};
struct NttCache {
    NttTables t[NTT_MAX_LOG + 1];
};

static NttCache& cache_of(hzk_ctx* ctx) {
    if (!ctx->ntt_cache) ctx->ntt_cache = new NttCache();
    return *(NttCache*)ctx->ntt_cache;
}

void ntt_release(hzk_ctx* ctx) {
    if (!ctx->ntt_cache) return;
    NttCache& c = *(NttCache*)ctx->ntt_cache;
    for (auto& t : c.t)
        for (auto p : {t.tw[0], t.tw[1], t.coset[0], t.coset[1], t.ninv})
            if (p) (void)hipFree(p);
    delete &c;
    ctx->ntt_cache = nullptr;
}

static int fill_powers(hzk_ctx* ctx, fr_t* out, size_t n, host::Fr w) {
    std::vector<host::Fr> sq(64);
    for (int b = 0; b < 64; ++b) { sq[b] = w; w = w.sqr(); }
    fr_t* dsq = nullptr;
    HZK_HIP(ctx, hipMalloc(&dsq, 64 * 32));
    HZK_HIP(ctx, hipMemcpy(dsq, sq.data(), 64 * 32, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_powers, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, out, n, (const fr_t*)dsq);
    HZK_HIP(ctx, hipGetLastError());
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HZK_HIP(ctx, hipFree(dsq));
    return 0;
}

static int ensure_tables(hzk_ctx* ctx, int logn, NttTables** out) {
    NttTables& t = cache_of(ctx).t[logn];
    *out = &t;
    if (t.ninv) return 0;
    size_t n = (size_t)1 << logn;
    size_t half = n / 2 ? n / 2 : 1;
    host::Fr w = root_of_unity(logn), wi = w.inverse();
    host::Fr g = fr_from_u64(5), gi = g.inverse();
    host::Fr ninv = fr_from_u64(n).inverse();
    for (int d = 0; d < 2; ++d) {
        HZK_HIP(ctx, hipMalloc(&t.tw[d], half * 32));
        HZK_HIP(ctx, hipMalloc(&t.coset[d], n * 32));
        int rc = fill_powers(ctx, t.tw[d], half, d ? wi : w);
        if (rc) return rc;
        rc = fill_powers(ctx, t.coset[d], n, d ? gi : g);
        if (rc) return rc;
    }
    fr_t* ni = nullptr;
    HZK_HIP(ctx, hipMalloc(&ni, 32));
    HZK_HIP(ctx, hipMemcpy(ni, ninv.v, 32, hipMemcpyHostToDevice));
    t.ninv = ni;
    return 0;
}

static void set_ntt_lds() {
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void*)k_ntt_pass, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (1 << NTT_TILE_BITS) * 32);
        (void)hipFuncSetAttribute((const void*)k_ntt_pass_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (1 << NTT_TILE_BITS) * 32);
    });
}

// In-place transform with twiddle table tw (w^e, e < n/2), optional device factor.
int ntt_run(hzk_ctx* ctx, fr_t* x, int k, const fr_t* tw, const fr_t* factor) {
    size_t n = (size_t)1 << k;
    hipStream_t st = ctx->stream;
    set_ntt_lds();
    if (k == 0) {
        if (factor) hipLaunchKernelGGL(k_bitrev_scale, dim3(1), dim3(1), 0, st, x, n, 1, factor);   // 1 x 1: scale only
        return 0;
    }
    // pass plan: the last pass takes the lowest min(k, 11) stages (contiguous
    // rows); the rest split evenly into passes of <= 9 stages (>= 4 columns)
    int last = k < NTT_TILE_BITS ? k : NTT_TILE_BITS;
    int rest = k - last;
    int npass = (rest + 8) / 9;
    int s0 = 0;
    {
        TimedLaunch tl(ctx, "ntt_pass");
        for (int p = 0; p < npass; ++p) {
            int R = rest / npass + (p < rest % npass ? 1 : 0);
            int gbits = NTT_TILE_BITS - R;
            int low = k - s0 - R;
            if (gbits > low) gbits = low;
            size_t tiles = n >> (R + gbits);
            hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)tiles), dim3(NTT_THREADS), (32u << (R + gbits)), st, x, tw, k, s0, R,
                               gbits);
            s0 += R;
        }
        hipLaunchKernelGGL(k_ntt_pass, dim3((unsigned)(n >> last)), dim3(NTT_THREADS), (32u << last), st, x, tw, k, s0, last, 0);
    }
    {
        TimedLaunch tl(ctx, "ntt_bitrev");
        hipLaunchKernelGGL(k_bitrev_scale, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, k, factor);
    }
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

static int log2_exact(size_t n) {
    int k = 0;
    while (((size_t)1 << k) < n) ++k;
    return ((size_t)1 << k) == n ? k : -1;
}

int ntt_standard(hzk_ctx* ctx, fr_t* x, size_t deg, int inverse) {
    int k = log2_exact(deg);
    HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG && k <= 28, HZK_ERR_INVALID, "ntt: deg %zu is not a power of two <= 2^%d", deg,
              NTT_MAX_LOG);
    NttTables* t;
    int rc = ensure_tables(ctx, k, &t);
    if (rc) return rc;
    return ntt_run(ctx, x, k, t->tw[inverse ? 1 : 0], inverse ? t->ninv : nullptr);
}

int coset_ntt(hzk_ctx* ctx, fr_t* x, size_t deg, int inverse) {
    int k = log2_exact(deg);
    HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "coset_ntt: deg %zu is not a power of two <= 2^%d", deg,
              NTT_MAX_LOG);
    NttTables* t;
    int rc = ensure_tables(ctx, k, &t);
    if (rc) return rc;
    if (!inverse) {
        {
            TimedLaunch tl(ctx, "ntt_coset_scale");
            hipLaunchKernelGGL(k_mul_table, dim3(blocks_for(deg, 256)), dim3(256), 0, ctx->stream, x, (const fr_t*)t->coset[0],
                               deg, (const fr_t*)nullptr);
        }
        return ntt_run(ctx, x, k, t->tw[0], nullptr);
    }
    rc = ntt_run(ctx, x, k, t->tw[1], t->ninv);
    if (rc) return rc;
    TimedLaunch tl(ctx, "ntt_coset_scale");
    hipLaunchKernelGGL(k_mul_table, dim3(blocks_for(deg, 256)), dim3(256), 0, ctx->stream, x, (const fr_t*)t->coset[1], deg,
                       (const fr_t*)nullptr);
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

// pass plan of ntt_run: (s0, R, gbits) per pass, the 11-stage contiguous pass last
static void plan_passes(int k, std::vector<int3>& out) {
    out.clear();
    if (k == 0) return;
    int last = k < NTT_TILE_BITS ? k : NTT_TILE_BITS;
    int rest = k - last;
    int npass = (rest + 8) / 9;
    int s0 = 0;
    for (int p = 0; p < npass; ++p) {
        int R = rest / npass + (p < rest % npass ? 1 : 0);
        int gbits = NTT_TILE_BITS - R;
        int low = k - s0 - R;
        if (gbits > low) gbits = low;
        out.push_back(make_int3(s0, R, gbits));
        s0 += R;
    }
    out.push_back(make_int3(s0, last, 0));
}

// Batched transforms: kind 0 NTT, 1 inverse NTT (scaled 1/n), 2 coset NTT,
// 3 coset inverse NTT. Passes are right-aligned into rounds so every
// transform's last (contiguous) pass runs in the final round; each round is one
// launch over the tiles of every transform in it, then one launch applies the
// bit-reverse permutation with the per-transform scale / coset factors.
int ntt_batch(hzk_ctx* ctx, const NttJob* jobs, size_t count, int capture) {
    HZK_CHECK(ctx, jobs && count > 0, HZK_ERR_INVALID, "ntt batch: empty");
    std::vector<int> ks(count);
    std::vector<NttTables*> tabs(count);
    std::vector<std::vector<int3>> plans(count);
    size_t rounds = 0;
    for (size_t i = 0; i < count; ++i) {
        HZK_CHECK(ctx, jobs[i].x && jobs[i].kind >= 0 && jobs[i].kind <= 3, HZK_ERR_INVALID, "ntt batch: bad job %zu", i);
        int k = log2_exact(jobs[i].deg);
        HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "ntt batch: deg %zu of job %zu is not a power of two <= 2^%d",
                  jobs[i].deg, i, NTT_MAX_LOG);
        ks[i] = k;
        int rc = ensure_tables(ctx, k, &tabs[i]);
        if (rc) return rc;
        plan_passes(k, plans[i]);
        rounds = std::max(rounds, plans[i].size());
    }
    set_ntt_lds();
    // pass descriptors, round by round
    std::vector<NttPassDesc> pds;
    std::vector<size_t> round_begin(rounds + 1, 0);
    std::vector<uint64_t> round_tiles(rounds, 0);
    std::vector<uint32_t> round_lds(rounds, 0);
    for (size_t r = 0; r < rounds; ++r) {
        round_begin[r] = pds.size();
        uint64_t tiles = 0;
        for (size_t i = 0; i < count; ++i) {
            size_t m = plans[i].size();
            if (m + r < rounds) continue;
            int3 pl = plans[i][m + r - rounds];
            int inv = jobs[i].kind & 1;
            NttPassDesc d{(fr_t*)jobs[i].x, tabs[i]->tw[inv], ks[i], pl.x, pl.y, pl.z, tiles};
            pds.push_back(d);
            tiles += (uint64_t)1 << (ks[i] - pl.y - pl.z);
            round_lds[r] = std::max(round_lds[r], 32u << (pl.y + pl.z));
        }
        round_tiles[r] = tiles;
    }
    round_begin[rounds] = pds.size();
    std::vector<NttElemDesc> eds(count);
    uint64_t elems = 0;
    bool any_pre = false;
    for (size_t i = 0; i < count; ++i) {
        int kind = jobs[i].kind;
        eds[i] = {(fr_t*)jobs[i].x, (kind & 1) ? tabs[i]->ninv : nullptr, kind == 2 ? tabs[i]->coset[0] : nullptr,
                  kind == 3 ? tabs[i]->coset[1] : nullptr, ks[i], elems};
        any_pre |= kind == 2;
        elems += jobs[i].deg;
    }
    // descriptors to the device
    size_t pbytes = pds.size() * sizeof(NttPassDesc), ebytes = eds.size() * sizeof(NttElemDesc);
    size_t need = ((pbytes + 255) & ~size_t(255)) + ebytes + 256;
    if (ctx->desc_bytes < need) {
        if (ctx->desc) { HZK_HIP(ctx, hipStreamSynchronize(ctx->stream)); HZK_HIP(ctx, hipFree(ctx->desc)); ctx->desc = nullptr; }
        HZK_HIP(ctx, hipMalloc(&ctx->desc, need * 2));
        ctx->desc_bytes = need * 2;
        ++ctx->desc_generation;
    }
    char* pin;
    int slot;
    int rc = stage_acquire(ctx, need, &pin, &slot);
    if (rc) return rc;
    size_t eoff = (pbytes + 255) & ~size_t(255);
    std::memcpy(pin, pds.data(), pbytes);
    std::memcpy(pin + eoff, eds.data(), ebytes);
    hipStream_t st = ctx->stream;
    HZK_HIP(ctx, hipMemcpyAsync(ctx->desc, pin, eoff + ebytes, hipMemcpyHostToDevice, st));
    rc = stage_fence(ctx, slot);
    if (rc) return rc;
    const NttPassDesc* dpd = (const NttPassDesc*)ctx->desc;
    const NttElemDesc* ded = (const NttElemDesc*)((char*)ctx->desc + eoff);

    auto stages = [&]() {
        if (any_pre)
            hipLaunchKernelGGL(k_ntt_pre_batch, dim3(blocks_for(elems, 256)), dim3(256), 0, st, ded, (uint32_t)count, elems);
        for (size_t r = 0; r < rounds; ++r) {
            uint32_t np = (uint32_t)(round_begin[r + 1] - round_begin[r]);
            hipLaunchKernelGGL(k_ntt_pass_batch, dim3((unsigned)round_tiles[r]), dim3(NTT_THREADS), round_lds[r], st,
                               dpd + round_begin[r], np);
        }
        hipLaunchKernelGGL(k_bitrev_batch, dim3(blocks_for(elems, 256)), dim3(256), 0, st, ded, (uint32_t)count, elems);
    };
    if (capture) {
        std::string key = "ntt:";
        for (size_t i = 0; i < count; ++i) key += std::to_string(ks[i]) + "/" + std::to_string(jobs[i].kind) + ";";
        hzk_graph_entry& g = ctx->graphs[key];
        if (!g.exec || g.generation != ctx->desc_generation) {
            if (g.exec) { (void)hipGraphExecDestroy(g.exec); g.exec = nullptr; }
            hipGraph_t graph;
            HZK_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            stages();
            hipError_t e = hipStreamEndCapture(st, &graph);
            HZK_HIP(ctx, e);
            e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HZK_HIP(ctx, e);
            g.generation = ctx->desc_generation;
        }
        TimedLaunch tl(ctx, "ntt_graph");
        HZK_HIP(ctx, hipGraphLaunch(g.exec, st));
    } else {
        TimedLaunch tl(ctx, "ntt_batch");
        stages();
    }
    HZK_HIP(ctx, hipGetLastError());
    return 0;
}

int ntt_standard_raw(hzk_ctx* ctx, void* x, size_t deg, int inverse) { return ntt_standard(ctx, (fr_t*)x, deg, inverse); }
int coset_ntt_raw(hzk_ctx* ctx, void* x, size_t deg, int inverse) { return coset_ntt(ctx, (fr_t*)x, deg, inverse); }

}  // namespace hzk

using namespace hzk;

extern "C" {

int hzk_ntt(hzk_ctx* ctx, void* coeffs, size_t deg, const void* omega_powers, int inverse, const void* factor) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, coeffs != nullptr, HZK_ERR_INVALID, "ntt: null coefficients");
    if (!omega_powers) {
        if (factor) {
            int k = log2_exact(deg);
            HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "ntt: deg %zu invalid", deg);
            NttTables* t;
            int rc = ensure_tables(ctx, k, &t);
            if (rc) return rc;
            return ntt_run(ctx, (fr_t*)coeffs, k, t->tw[inverse ? 1 : 0], (const fr_t*)factor);
        }
        return ntt_standard(ctx, (fr_t*)coeffs, deg, inverse);
    }
    int k = log2_exact(deg);
    HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "ntt: deg %zu is not a power of two <= 2^%d", deg, NTT_MAX_LOG);
    return ntt_run(ctx, (fr_t*)coeffs, k, (const fr_t*)omega_powers, (const fr_t*)factor);
}

int hzk_ntt_batch(hzk_ctx* ctx, const hzk_ntt_desc* d, size_t count, uint32_t flags) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, d && count > 0, HZK_ERR_INVALID, "ntt batch: empty");
    HZK_CHECK(ctx, (flags & ~HZK_BATCH_GRAPH) == 0, HZK_ERR_INVALID, "ntt batch: bad flags %u", flags);
    std::vector<NttJob> jobs(count);
    for (size_t i = 0; i < count; ++i) jobs[i] = {d[i].coeffs, d[i].deg, d[i].kind};
    return ntt_batch(ctx, jobs.data(), count, (flags & HZK_BATCH_GRAPH) ? 1 : 0);
}

// The reference's twiddle table for an M-point interpolation domain and its
// 2M-point evaluation domain (ntt_setup, include/protocol/groth16-operator.cuh:
// 6-32), on the host: w_M^i and w_M^-i (i < M), M^-1, w_2M^i and w_2M^-i
// (i < 2M), then (4M)^-1 at 6M + 1, the entry the reference writes there
// (its layout comment, groth16.cuh:32-33, names (2M)^-1).
int hzk_ntt_setup(size_t M, void* host_out) {
    int lg = log2_exact(M);
    if (lg < 0 || lg + 2 > 28 || !host_out) return HZK_ERR_INVALID;
    host::Fr* t = (host::Fr*)host_out;
    const host::Fr w[4] = {root_of_unity(lg), root_of_unity(lg).inverse(), root_of_unity(lg + 1),
                           root_of_unity(lg + 1).inverse()};
    const size_t base[4] = {0, M, 2 * M + 1, 4 * M + 1}, len[4] = {M, M, 2 * M, 2 * M};
    const size_t chunk = 4096;
    for (int s = 0; s < 4; ++s) {
        const long nchunks = (long)((len[s] + chunk - 1) / chunk);
#pragma omp parallel for schedule(static)
        for (long c = 0; c < nchunks; ++c) {
            size_t i0 = (size_t)c * chunk, i1 = std::min(len[s], i0 + chunk);
            uint64_t e[4] = {(uint64_t)i0, 0, 0, 0};
            host::Fr acc = w[s].pow(e);
            for (size_t i = i0; i < i1; ++i) {
                t[base[s] + i] = acc;
                acc = acc * w[s];
            }
        }
    }
    host::Fr two_inv = fr_from_u64(2).inverse();
    host::Fr mi = host::Fr::one(), m4i;
    for (int i = 0; i < lg; ++i) mi = mi * two_inv;
    m4i = mi * two_inv * two_inv;
    t[2 * M] = mi;
    t[6 * M + 1] = m4i;
    return 0;
}

int hzk_coset_ntt(hzk_ctx* ctx, void* coeffs, size_t deg, int inverse) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, coeffs != nullptr, HZK_ERR_INVALID, "coset_ntt: null coefficients");
    return coset_ntt(ctx, (fr_t*)coeffs, deg, inverse);
}

}  // extern "C"
