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
// Here the log2(deg) stages are grouped into passes. A pass covers R
// consecutive stages over tiles of 2^R rows x 2^gbits columns (2^11 elements,
// 64 KiB); columns are contiguous in HBM, so loads and stores stay coalesced.
// Inside a tile every thread holds 8 elements in registers: the 2^3 rows
// that three consecutive stages pair up (a radix-8 group, or a radix-4 /
// radix-2 group padded with a neighbouring bit), runs those stages' 12
// butterflies there with interleaved products, and trades its elements with
// the other threads through LDS before the next three stages: 4 LDS
// exchanges for 11 stages instead of 11 LDS round trips and barriers. 2^20
// takes two passes (9 stages, then the last 11, whose rows are contiguous).
//
// Decimation in frequency (DIF, the reference's butterfly) takes natural
// order to bit-reversed order; running the same stages backwards with the
// transposed butterfly (a, b) -> (a + t b, a - t b) (decimation in time,
// DIT) takes bit-reversed order back to natural order. The operator keeps
// the reference's natural-in, natural-out contract with one more pass that
// applies the bit-reverse permutation fused with the scale factor. The
// Groth16 H pipeline needs no permutation at all: DIF for its inverse
// transforms, DIT for its coset transform, with the coset powers and 1/n
// folded into the first load or the last store of a pass, and the quotient
// h left in bit-reversed order, matched by a bit-reversed copy of the H
// query points.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>
#include "common.hpp"
#include "field.hpp"
#include "fr29.hpp"
#include "host_ec.hpp"

namespace hzk {

// Tiles of 2^TB elements, TB = 9, 10 or 11, 2^(TB-LE) threads each holding
// 2^LE elements: LE = 3 (a radix-8 group) where that gives every SIMD two
// waves, else LE = 1 (launches too small for that, e.g. a lone 2^18
// transform: two-element threads, 2048 waves). An LE = 3 launch takes the
// largest tile that still gives every CU two workgroups (batches and 2^20 and
// up: 2^11); an LE = 1 launch the largest that gives every CU one (2^18: 2^10).
static constexpr int NTT_TILE_BITS = 11;   // largest tile
template <int TB, int LE = 3>
constexpr int ntt_threads() { return (1 << TB) >> LE; }
static constexpr int NTT_MAX_LOG = 26;

// one pass of one transform over its tiles [tile_begin, next tile_begin)
struct NttPassDesc {
    fr_t* x;
    const fr_t* tw;     // w^e (or w^-e), e < n/2
    const fr_t* pre;    // x[j] *= pre[j] as the pass loads (first pass only), or null
    const fr_t* post;   // x[j] *= post[j] as the pass stores (last pass only), or null
    int k, s0, R, gbits, dit;
    uint64_t tile_begin;
};

// LDS slot of tile element e: two 16-byte planes (low and high half of the
// element), so lanes reading consecutive elements touch consecutive 16-byte
// words; the low 3 bits are XOR-ed with the next 3 so that threads holding
// 8 consecutive elements (the lowest radix-8 group) spread over the banks
HZK_DEV uint32_t lds_slot(uint32_t e) { return e ^ ((e >> 3) & 7); }
template <int TB>
HZK_DEV void lds_put(uint4* sh, uint32_t e, const fr_t& v) {
    const uint32_t s = lds_slot(e);
    sh[s] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
    sh[s + (1u << TB)] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
}
template <int TB>
HZK_DEV fr_t lds_get(const uint4* sh, uint32_t e) {
    const uint32_t s = lds_slot(e);
    const uint4 a = sh[s], b = sh[s + (1u << TB)];
    fr_t v;
    v.l[0] = a.x, v.l[1] = a.y, v.l[2] = a.z, v.l[3] = a.w, v.l[4] = b.x, v.l[5] = b.y, v.l[6] = b.z, v.l[7] = b.w;
    return v;
}

// Geometry of one tile of a pass (stages s0 .. s0+R-1 of a 2^k transform).
// Tile element e has row m = e >> gbits (the butterfly dimension) and column
// g = e & (2^gbits - 1); its global index is base + (m << low) + g. The DIF
// stage s0 + st pairs rows differing in bit R - 1 - st, i.e. tile bit
// gbits + R - 1 - st, with twiddle w^((((m & mask) << low) | column) << s).
struct NttTile {
    int k, s0, R, gbits, low, TB;
    uint32_t colbase;
    size_t base;
    HZK_DEV NttTile(int k_, int s0_, int R_, int gbits_, uint32_t b) : k(k_), s0(s0_), R(R_), gbits(gbits_) {
        low = k - s0 - R;
        TB = R + gbits;
        const uint32_t lchunks = (1u << low) >> gbits;
        const uint32_t hi = b / lchunks, lc = b % lchunks;
        base = ((size_t)hi << (k - s0)) | ((size_t)lc << gbits);
        colbase = lc << gbits;
    }
    HZK_DEV size_t gidx(uint32_t e) const { return base + ((size_t)(e >> gbits) << low) + (e & ((1u << gbits) - 1)); }
    // twiddle exponent of the butterfly of local stage st whose lower element is e
    HZK_DEV size_t texp(uint32_t e, int st) const {
        const int mp = R - 1 - st;
        const uint32_t m = (e >> gbits) & ((1u << mp) - 1);
        return ((((size_t)m) << low) | colbase | (e & ((1u << gbits) - 1))) << (s0 + st);
    }
};

// the 2^(LE-1) butterflies of one stage over the thread's 2^LE elements v[i]
// (element i at tile index eb | (i << L)); the stage pairs bit IB of i
template <int IB, bool DIT, int LE>
HZK_DEV void ntt_stage(fr_t* v, const NttTile& T, const fr_t* __restrict__ tw, uint32_t eb, int L, int st) {
    constexpr int H = 1 << (LE - 1);   // butterflies
    int lo[H], hi[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const int below = j & ((1 << IB) - 1), above = (j >> IB) << (IB + 1);
        lo[j] = above | below;
        hi[j] = lo[j] | (1 << IB);
    }
    fr_t t[H];
#pragma unroll
    for (int j = 0; j < H; ++j) t[j] = tw[T.texp(eb | ((uint32_t)lo[j] << L), st)];
    if (!DIT) {
        fr_t d[H];
#pragma unroll
        for (int j = 0; j < H; ++j) {
            d[j] = v[lo[j]] - v[hi[j]];
            v[lo[j]] = v[lo[j]] + v[hi[j]];
        }
        if constexpr (H == 1) v[hi[0]] = d[0] * t[0];
        else
#pragma unroll
            for (int j = 0; j < H; j += 2) fr_t::mul2(v[hi[j]], d[j], t[j], v[hi[j + 1]], d[j + 1], t[j + 1]);
    } else {
        fr_t p[H];
        if constexpr (H == 1) p[0] = v[hi[0]] * t[0];
        else
#pragma unroll
            for (int j = 0; j < H; j += 2) fr_t::mul2(p[j], v[hi[j]], t[j], p[j + 1], v[hi[j + 1]], t[j + 1]);
#pragma unroll
        for (int j = 0; j < H; ++j) {
            v[hi[j]] = v[lo[j]] - p[j];
            v[lo[j]] = v[lo[j]] + p[j];
        }
    }
}

template <bool DIT, int LE>
HZK_DEV void ntt_stage_rt(fr_t* v, const NttTile& T, const fr_t* __restrict__ tw, uint32_t eb, int L, int st, int ib) {
    if constexpr (LE == 1) ntt_stage<0, DIT, 1>(v, T, tw, eb, L, st);
    else if constexpr (LE == 2) {
        if (ib == 0) ntt_stage<0, DIT, 2>(v, T, tw, eb, L, st);
        else ntt_stage<1, DIT, 2>(v, T, tw, eb, L, st);
    } else {
        if (ib == 0) ntt_stage<0, DIT, 3>(v, T, tw, eb, L, st);
        else if (ib == 1) ntt_stage<1, DIT, 3>(v, T, tw, eb, L, st);
        else ntt_stage<2, DIT, 3>(v, T, tw, eb, L, st);
    }
}

// x[i] *= tab[i] for the thread's 2^LE elements, interleaved in pairs
template <int LE>
HZK_DEV void mul_table(fr_t* v, const NttTile& T, const fr_t* __restrict__ tab, uint32_t eb, int L) {
#pragma unroll
    for (int i = 0; i < (1 << LE); i += 2) {
        const fr_t a = tab[T.gidx(eb | ((uint32_t)i << L))], b = tab[T.gidx(eb | ((uint32_t)(i + 1) << L))];
        fr_t::mul2(v[i], v[i], a, v[i + 1], v[i + 1], b);
    }
}

// One tile of a pass: stages in chunks of up to LE. DIF runs local stages
// 0..R-1 from the first chunk (tile bits TB-1..) down; DIT runs them in
// reverse. A thread's 2^LE elements for a chunk span the LE tile bits
// [L, L + LE) that contain the chunk's stage bits; its other tile bits come
// from the thread index (low bits first, so a wave's lanes read consecutive
// elements wherever L > 0).
template <bool DIT, int TB, int LE = 3>
HZK_DEV void ntt_tile(fr_t* __restrict__ x, const fr_t* __restrict__ tw, const fr_t* __restrict__ pre,
                      const fr_t* __restrict__ post, int k, int s0, int R, int gbits, uint32_t b, uint4* sh) {
    constexpr int E = 1 << LE;
    const NttTile T(k, s0, R, gbits, b);
    const uint32_t t = threadIdx.x;
    const bool active = t < (1u << (T.TB - LE));
    fr_t v[E];
    const int nch = (R + LE - 1) / LE;
    uint32_t eb_prev = 0;
    int L_prev = 0;
    for (int c = 0; c < nch; ++c) {
        int stA, len;
        if (!DIT) {
            stA = LE * c;
            len = min(LE, R - stA);
        } else {
            stA = max(0, R - LE * (c + 1));
            len = R - LE * c - stA;
        }
        const int btop = gbits + R - 1 - stA, blo = btop - len + 1;
        const int L = min(blo, T.TB - LE);
        const uint32_t eb = (t & ((1u << L) - 1)) | ((t >> L) << (L + LE));
        if (c == 0) {
            if (active) {
#pragma unroll
                for (int i = 0; i < E; ++i) v[i] = x[T.gidx(eb | ((uint32_t)i << L))];
                if (pre) mul_table<LE>(v, T, pre, eb, L);
            }
        } else {
            if (c > 1) __syncthreads();   // every thread has read the previous exchange
            if (active)
#pragma unroll
                for (int i = 0; i < E; ++i) lds_put<TB>(sh, eb_prev | ((uint32_t)i << L_prev), v[i]);
            __syncthreads();
            if (active)
#pragma unroll
                for (int i = 0; i < E; ++i) v[i] = lds_get<TB>(sh, eb | ((uint32_t)i << L));
        }
        if (active) {
            if (!DIT)
                for (int st = stA; st < stA + len; ++st) ntt_stage_rt<false, LE>(v, T, tw, eb, L, st, gbits + R - 1 - st - L);
            else
                for (int st = stA + len - 1; st >= stA; --st) ntt_stage_rt<true, LE>(v, T, tw, eb, L, st, gbits + R - 1 - st - L);
        }
        eb_prev = eb;
        L_prev = L;
    }
    if (active) {
        if (post) mul_table<LE>(v, T, post, eb_prev, L_prev);
#pragma unroll
        for (int i = 0; i < E; ++i) x[T.gidx(eb_prev | ((uint32_t)i << L_prev))] = v[i];
    }
}

// The same pass with two elements per thread (one butterfly per stage, an
// LDS exchange between stages) in fr29 arithmetic, for lone transforms that
// are latency-bound: each stage's twiddle is loaded one stage ahead, so no
// butterfly waits for a global load.
template <bool DIT, int TB>
HZK_DEV void ntt_tile_r2(fr_t* __restrict__ x, const fr_t* __restrict__ tw, const fr_t* __restrict__ pre,
                         const fr_t* __restrict__ post, int k, int s0, int R, int gbits, uint32_t b, uint4* sh) {
    const NttTile T(k, s0, R, gbits, b);
    const uint32_t t = threadIdx.x;
    const bool active = t < (1u << (T.TB - 1));
    // executed step i runs local stage st(i) on tile bit L(i) = gbits + R - 1 - st(i);
    // the thread's pair is (eb(i), eb(i) | 2^L(i))
    auto stage_of = [&](int i) { return DIT ? R - 1 - i : i; };
    auto bit_of = [&](int i) { return gbits + R - 1 - stage_of(i); };
    auto pair_base = [&](int L) { return (t & ((1u << L) - 1)) | ((t >> L) << (L + 1)); };
    fr29 v0, v1;
    fr_t w, wn;
    uint32_t eb = 0;
    int L = 0;
    bool lds_used = false;
    if (active) w = tw[T.texp(pair_base(bit_of(0)), stage_of(0))];
    for (int i = 0; i < R; ++i) {
        const int Li = bit_of(i);
        const uint32_t ebi = pair_base(Li);
        if (active && i + 1 < R) wn = tw[T.texp(pair_base(bit_of(i + 1)), stage_of(i + 1))];
        if (i == 0) {
            if (active) {
                v0 = fr29::unpack(x[T.gidx(ebi)]).reduce();
                v1 = fr29::unpack(x[T.gidx(ebi | (1u << Li))]).reduce();
                if (pre) {
                    v0 = fr29::mul(v0, fr29::unpack(pre[T.gidx(ebi)]).times32());
                    v1 = fr29::mul(v1, fr29::unpack(pre[T.gidx(ebi | (1u << Li))]).times32());
                }
            }
        } else if (min(L, Li) < 6) {
            // the pair bits of this stage and the last differ by one; the
            // threads that trade elements differ in thread bit pb = min(L, Li)
            // and share a wave: swap one element across lanes, no barrier.
            // Thread bit pb = b keeps v_b and receives the partner's element
            // into v_(1-b) (both for DIF, pb = Li, and DIT, pb = L).
            const int pb = min(L, Li);
            const bool hi = (t >> pb) & 1;
            fr29 send = hi ? v0 : v1, got;
#pragma unroll
            for (int q = 0; q < 9; ++q) got.l[q] = (uint32_t)__shfl_xor((int)send.l[q], 1 << pb, 64);
            if (hi) v0 = got;
            else v1 = got;
        } else {
            if (lds_used) __syncthreads();   // every thread has read the previous LDS exchange
            if (active) {
                lds_put<TB>(sh, eb, v0.pack());
                lds_put<TB>(sh, eb | (1u << L), v1.pack());
            }
            __syncthreads();
            if (active) {
                v0 = fr29::unpack(lds_get<TB>(sh, ebi));
                v1 = fr29::unpack(lds_get<TB>(sh, ebi | (1u << Li)));
            }
            lds_used = true;
        }
        if (active) {
            const fr29 w32 = fr29::unpack(w).times32();   // canonical twiddle: < 32r
            if (!DIT) {
                const fr29 d = fr29::sub2(v0, v1);         // < 4r
                v0 = fr29::add(v0, v1).reduce();
                v1 = fr29::mul(d, w32);                    // (4)(32) < 169: < 2r
            } else {
                const fr29 p = fr29::mul(v1, w32);         // (2)(32)
                v1 = fr29::sub2(v0, p).reduce();
                v0 = fr29::add(v0, p).reduce();
            }
        }
        w = wn;
        eb = ebi;
        L = Li;
    }
    if (active) {
        if (post) {
            v0 = fr29::mul(v0, fr29::unpack(post[T.gidx(eb)]).times32());
            v1 = fr29::mul(v1, fr29::unpack(post[T.gidx(eb | (1u << L))]).times32());
        }
        x[T.gidx(eb)] = v0.canonical().pack();
        x[T.gidx(eb | (1u << L))] = v1.canonical().pack();
    }
}

template <bool DIT, int TB, int LE>
__global__ void __launch_bounds__((ntt_threads<TB, LE>())) k_ntt_pass(fr_t* __restrict__ x, const fr_t* __restrict__ tw,
                                                                    const fr_t* __restrict__ pre, const fr_t* __restrict__ post,
                                                                    int k, int s0, int R, int gbits) {
    __shared__ uint4 sh[2 << TB];
    if constexpr (LE == 1) ntt_tile_r2<DIT, TB>(x, tw, pre, post, k, s0, R, gbits, blockIdx.x, sh);
    else ntt_tile<DIT, TB, LE>(x, tw, pre, post, k, s0, R, gbits, blockIdx.x, sh);
}

// transforms of 2 and 4 points (too small for an 8-element group): one thread each
__global__ void k_ntt_tiny(fr_t* __restrict__ x, const fr_t* __restrict__ tw, int k) {
    if (threadIdx.x != 0) return;
    const int n = 1 << k;
    for (int s = 0; s < k; ++s) {
        const int h = n >> (s + 1);
        for (int j = 0; j < n; ++j) {
            if (j & h) continue;
            const fr_t a = x[j], c = x[j + h];
            x[j] = a + c;
            x[j + h] = (a - c) * tw[(j & (h - 1)) << s];
        }
    }
}

// ------------------------------------------------------------------ batched transforms
// one transform for the element-wise stages: its elements occupy [begin, next begin)
struct NttElemDesc {
    fr_t* x;
    const fr_t* factor;   // bit-reverse scale (1/n for inverses) or null
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

template <int TB, int LE>
__global__ void __launch_bounds__((ntt_threads<TB, LE>())) k_ntt_pass_batch(const NttPassDesc* __restrict__ pd, uint32_t npd) {
    __shared__ uint4 sh[2 << TB];
    const uint32_t i = find_desc(pd, npd, blockIdx.x);
    const NttPassDesc& d = pd[i];
    const uint32_t b = (uint32_t)(blockIdx.x - d.tile_begin);
    if constexpr (LE == 1) {
        if (d.dit) ntt_tile_r2<true, TB>(d.x, d.tw, d.pre, d.post, d.k, d.s0, d.R, d.gbits, b, sh);
        else ntt_tile_r2<false, TB>(d.x, d.tw, d.pre, d.post, d.k, d.s0, d.R, d.gbits, b, sh);
    } else {
        if (d.dit) ntt_tile<true, TB, LE>(d.x, d.tw, d.pre, d.post, d.k, d.s0, d.R, d.gbits, b, sh);
        else ntt_tile<false, TB, LE>(d.x, d.tw, d.pre, d.post, d.k, d.s0, d.R, d.gbits, b, sh);
    }
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
// The same for k >= 8 through LDS in 16 x 16 blocks: index i = (a, m, b)
// with a its top 4 bits, b its low 4 and m the k - 8 bits between, so
// rev(i) = (rev b, rev m, rev a): block m (every a, b) and block rev(m) trade
// places, transposed, and every global access is a run of 16 consecutive
// elements. One workgroup per m; the one with m > rev(m) has nothing to do.
// factor scales every element; post[p] the one landing at p.
HZK_DEV void bitrev_tile(fr_t* __restrict__ x, int k, uint32_t m, const fr_t* __restrict__ factor,
                         const fr_t* __restrict__ post, fr_t (*blk)[16 * 17]) {
    const int mb = k - 8;
    const uint32_t mr = mb ? __brev(m) >> (32 - mb) : 0;
    if (m > mr) return;
    const uint32_t t = threadIdx.x, a = t >> 4, b = t & 15;
    const size_t stride_a = (size_t)1 << (k - 4);
    auto at = [&](uint32_t aa, uint32_t mm, uint32_t bb) { return (size_t)aa * stride_a + ((size_t)mm << 4) + bb; };
    fr_t v0 = x[at(a, m, b)], v1 = x[at(a, mr, b)];
    if (factor) {
        const fr_t f = *factor;
        fr_t::mul2(v0, v0, f, v1, v1, f);
    }
    blk[0][a * 17 + b] = v0;   // rows padded to 17 elements
    blk[1][a * 17 + b] = v1;
    __syncthreads();
    // destination (a, mr, b) takes source (rev b, m, rev a), and (a, m, b) takes (rev b, mr, rev a)
    const uint32_t src = (__brev(b) >> 28) * 17 + (__brev(a) >> 28);
    const size_t d0 = at(a, mr, b), d1 = at(a, m, b);
    v0 = blk[0][src];
    v1 = blk[1][src];
    if (post) fr_t::mul2(v0, v0, post[d0], v1, v1, post[d1]);
    x[d0] = v0;
    if (m != mr) x[d1] = v1;
}
__global__ void __launch_bounds__(256) k_bitrev_tiles(fr_t* __restrict__ x, int k, const fr_t* __restrict__ factor) {
    __shared__ fr_t blk[2][16 * 17];
    bitrev_tile(x, k, blockIdx.x, factor, nullptr, blk);
}
// batched: transforms of 2^8 points and more, begin counting blocks (2^(k-8) each)
__global__ void __launch_bounds__(256) k_bitrev_tiles_batch(const NttElemDesc* __restrict__ ed, uint32_t ne) {
    __shared__ fr_t blk[2][16 * 17];
    const NttElemDesc& d = ed[find_elem(ed, ne, blockIdx.x)];
    bitrev_tile(d.x, d.k, (uint32_t)(blockIdx.x - d.begin), d.factor, d.post, blk);
}

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

// out[p] = tab[rev_k(p)] * scale: a bit-reversed copy of a coset table
__global__ void __launch_bounds__(256) k_bitrev_table(fr_t* __restrict__ out, const fr_t* __restrict__ tab, int k,
                                                      const fr_t* __restrict__ scale) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ((size_t)1 << k)) return;
    size_t r = k ? (size_t)(__brev((uint32_t)p) >> (32 - k)) : 0;
    out[p] = tab[r] * *scale;
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
// g^j and g^-j (n each), 1/n, and (built when the Groth16 H pipeline first
// asks) the bit-reversed coset tables g^rev(p) / n and g^-rev(p) / n; built
// on first use, kept for the context
struct NttTables {
    fr_t* tw[2] = {nullptr, nullptr};
    fr_t* coset[2] = {nullptr, nullptr};
    fr_t* coset_rev[2] = {nullptr, nullptr};
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
        for (auto p : {t.tw[0], t.tw[1], t.coset[0], t.coset[1], t.coset_rev[0], t.coset_rev[1], t.ninv})
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

static int ensure_rev_tables(hzk_ctx* ctx, int logn, NttTables* t) {
    if (t->coset_rev[0]) return 0;
    const size_t n = (size_t)1 << logn;
    for (int d = 0; d < 2; ++d) {
        HZK_HIP(ctx, hipMalloc(&t->coset_rev[d], n * 32));
        hipLaunchKernelGGL(k_bitrev_table, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, t->coset_rev[d],
                           (const fr_t*)t->coset[d], logn, (const fr_t*)t->ninv);
        HZK_HIP(ctx, hipGetLastError());
    }
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

static int log2_exact(size_t n) {
    int k = 0;
    while (((size_t)1 << k) < n) ++k;
    return ((size_t)1 << k) == n ? k : -1;
}

// Pass plan of a 2^k transform in DIF order over tiles of 2^TB elements,
// (s0, R, gbits) per pass: the lowest min(k, TB) stages (contiguous rows)
// form the last pass, the rest split evenly into passes of <= TB - 2 stages
// (>= 4 columns per tile). A DIT run executes the same passes in reverse
// order. k >= 3.
static void plan_passes(int k, int TB, std::vector<int3>& out) {
    out.clear();
    int last = k < TB ? k : TB;
    int rest = k - last;
    int npass = (rest + TB - 3) / (TB - 2);
    int s0 = 0;
    for (int p = 0; p < npass; ++p) {
        int R = rest / npass + (p < rest % npass ? 1 : 0);
        int gbits = TB - R;
        int low = k - s0 - R;
        if (gbits > low) gbits = low;
        out.push_back(make_int3(s0, R, gbits));
        s0 += R;
    }
    out.push_back(make_int3(s0, last, 0));
}

// The passes of one transform on ctx's stream: DIF (natural -> bit-reversed)
// or DIT (bit-reversed -> natural), pre applied by the first pass executed,
// post by the last. k >= 3.
// tile bits for a launch over `elems` elements: the largest tile that still
// gives every CU two workgroups, within 9..11
static int tile_bits_for(hzk_ctx* ctx, uint64_t elems) {
    int TB = NTT_TILE_BITS;
    while (TB > 9 && (elems >> TB) < 2 * (uint64_t)ctx->num_cus) --TB;
    return TB;
}

// elements per thread (2^LE) of a launch over `elems` elements: 8 while
// that still gives every SIMD two waves, else 2 (more threads and more
// exchanges: such transforms are latency-bound, not issue-bound; 4 measured
// slower than 2)
static int elem_bits_for(hzk_ctx* ctx, uint64_t elems) {
    const uint64_t want = 2ull * 4 * 64 * (uint64_t)ctx->num_cus;   // threads for two waves per SIMD
    return (elems >> 3) >= want ? 3 : 1;
}
// tile bits of a launch with 2^LE elements per thread: the largest tile that
// gives every CU two workgroups (LE = 3) or one workgroup of 2^(TB-1)
// threads (LE = 1: already two waves per SIMD, in fewer passes)
static int tile_bits_le(hzk_ctx* ctx, uint64_t elems, int le) {
    int TB = tile_bits_for(ctx, elems);
    if (le == 1)
        while (TB < NTT_TILE_BITS && (elems >> (TB + 1)) >= (uint64_t)ctx->num_cus) ++TB;
    return TB;
}

template <bool DIT, int TB, int LE>
static void launch_pass(hipStream_t st, unsigned tiles, fr_t* x, const fr_t* tw, const fr_t* pp, const fr_t* qq, int k, int3 pl) {
    hipLaunchKernelGGL((k_ntt_pass<DIT, TB, LE>), dim3(tiles), dim3(ntt_threads<TB, LE>()), 0, st, x, tw, pp, qq, k, pl.x, pl.y,
                       pl.z);
}
template <bool DIT, int TB>
static void launch_pass_le(int le, hipStream_t st, unsigned tiles, fr_t* x, const fr_t* tw, const fr_t* pp, const fr_t* qq, int k,
                           int3 pl) {
    if (le == 1) launch_pass<DIT, TB, 1>(st, tiles, x, tw, pp, qq, k, pl);
    else launch_pass<DIT, TB, 3>(st, tiles, x, tw, pp, qq, k, pl);
}

static void ntt_passes(hzk_ctx* ctx, fr_t* x, int k, const fr_t* tw, int dit, const fr_t* pre, const fr_t* post) {
    const int le = elem_bits_for(ctx, (uint64_t)1 << k);
    const int TB = std::min(k, tile_bits_le(ctx, (uint64_t)1 << k, le));
    std::vector<int3> plan;
    plan_passes(k, std::max(TB, 3), plan);
    if (dit) std::reverse(plan.begin(), plan.end());
    for (size_t p = 0; p < plan.size(); ++p) {
        const int3 pl = plan[p];
        const unsigned tiles = (unsigned)(((size_t)1 << k) >> (pl.y + pl.z));
        const fr_t* pp = p == 0 ? pre : nullptr;
        const fr_t* qq = p + 1 == plan.size() ? post : nullptr;
        // the kernel's tile holds 2^(R + gbits) <= 2^TB elements
        const int tb = pl.y + pl.z;
        if (tb <= 9) {
            if (dit) launch_pass_le<true, 9>(le, ctx->stream, tiles, x, tw, pp, qq, k, pl);
            else launch_pass_le<false, 9>(le, ctx->stream, tiles, x, tw, pp, qq, k, pl);
        } else if (tb == 10) {
            if (dit) launch_pass_le<true, 10>(le, ctx->stream, tiles, x, tw, pp, qq, k, pl);
            else launch_pass_le<false, 10>(le, ctx->stream, tiles, x, tw, pp, qq, k, pl);
        } else {
            if (dit) launch_pass_le<true, 11>(le, ctx->stream, tiles, x, tw, pp, qq, k, pl);
            else launch_pass_le<false, 11>(le, ctx->stream, tiles, x, tw, pp, qq, k, pl);
        }
    }
}

// In-place transform with twiddle table tw (w^e, e < n/2) and the
// reference's natural-in, natural-out contract; optional device factor.
int ntt_run(hzk_ctx* ctx, fr_t* x, int k, const fr_t* tw, const fr_t* factor) {
    size_t n = (size_t)1 << k;
    hipStream_t st = ctx->stream;
    if (k == 0) {
        if (factor) hipLaunchKernelGGL(k_bitrev_scale, dim3(1), dim3(1), 0, st, x, n, 1, factor);   // 1 x 1: scale only
        HZK_HIP(ctx, hipGetLastError());
        return 0;
    }
    {
        TimedLaunch tl(ctx, "ntt_pass");
        if (k < 3) hipLaunchKernelGGL(k_ntt_tiny, dim3(1), dim3(64), 0, st, x, tw, k);
        else ntt_passes(ctx, x, k, tw, 0, nullptr, nullptr);
    }
    {
        TimedLaunch tl(ctx, "ntt_bitrev");
        if (k >= 8) hipLaunchKernelGGL(k_bitrev_tiles, dim3(1u << (k - 8)), dim3(256), 0, st, x, k, factor);
        else hipLaunchKernelGGL(k_bitrev_scale, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, n, k, factor);
    }
    HZK_HIP(ctx, hipGetLastError());
    return 0;
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
        if (k >= 3) {   // the coset powers ride on the first pass's loads
            {
                TimedLaunch tl(ctx, "ntt_pass");
                ntt_passes(ctx, x, k, t->tw[0], 0, t->coset[0], nullptr);
            }
            TimedLaunch tl(ctx, "ntt_bitrev");
            if (k >= 8)
                hipLaunchKernelGGL(k_bitrev_tiles, dim3(1u << (k - 8)), dim3(256), 0, ctx->stream, x, k, (const fr_t*)nullptr);
            else
                hipLaunchKernelGGL(k_bitrev_scale, dim3(blocks_for(deg, 256)), dim3(256), 0, ctx->stream, x, deg, k,
                                   (const fr_t*)nullptr);
            HZK_HIP(ctx, hipGetLastError());
            return 0;
        }
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

// Batched transforms. kind 0 NTT, 1 inverse NTT (scaled 1/n), 2 coset NTT,
// 3 coset inverse NTT, all natural order in and out; and, for the Groth16
// H pipeline (no permutation): 4 inverse NTT, natural in, bit-reversed out,
// unscaled; 5 coset NTT of bit-reversed coefficients c scaled by 1/n on the
// way in (x[p] * g^rev(p) / n), natural out; 6 coset inverse NTT, natural
// in, bit-reversed out, with g^-rev(p) / n applied on the way out. Pass p of
// every transform runs in round p, each round one launch over the tiles of
// every transform in it; then one launch applies the bit-reverse
// permutation with the per-transform scale / coset factors to kinds 0..3.
int ntt_batch(hzk_ctx* ctx, const NttJob* jobs, size_t count, int capture) {
    HZK_CHECK(ctx, jobs && count > 0, HZK_ERR_INVALID, "ntt batch: empty");
    std::vector<int> ks(count);
    std::vector<NttTables*> tabs(count);
    std::vector<std::vector<int3>> plans(count);
    size_t rounds = 0;
    uint64_t all = 0;
    for (size_t i = 0; i < count; ++i) all += jobs[i].deg;
    const int LE = elem_bits_for(ctx, all);
    const int TB = tile_bits_le(ctx, all, LE);
    for (size_t i = 0; i < count; ++i) {
        HZK_CHECK(ctx, jobs[i].x && jobs[i].kind >= 0 && jobs[i].kind <= 6, HZK_ERR_INVALID, "ntt batch: bad job %zu", i);
        int k = log2_exact(jobs[i].deg);
        HZK_CHECK(ctx, k >= 0 && k <= NTT_MAX_LOG, HZK_ERR_INVALID, "ntt batch: deg %zu of job %zu is not a power of two <= 2^%d",
                  jobs[i].deg, i, NTT_MAX_LOG);
        HZK_CHECK(ctx, jobs[i].kind <= 3 || k >= 3, HZK_ERR_INVALID, "ntt batch: kind %d needs deg >= 8", jobs[i].kind);
        ks[i] = k;
        int rc = ensure_tables(ctx, k, &tabs[i]);
        if (!rc && jobs[i].kind >= 5) rc = ensure_rev_tables(ctx, k, tabs[i]);
        if (rc) return rc;
        if (k >= 3) {
            plan_passes(k, TB, plans[i]);
            if (jobs[i].kind == 5) std::reverse(plans[i].begin(), plans[i].end());
        }
        rounds = std::max(rounds, plans[i].size());
    }
    // pass descriptors, round by round
    std::vector<NttPassDesc> pds;
    std::vector<size_t> round_begin(rounds + 1, 0);
    std::vector<uint64_t> round_tiles(rounds, 0);
    for (size_t r = 0; r < rounds; ++r) {
        round_begin[r] = pds.size();
        uint64_t tiles = 0;
        for (size_t i = 0; i < count; ++i) {
            const size_t m = plans[i].size();
            if (r >= m) continue;
            const int3 pl = plans[i][r];
            const int kind = jobs[i].kind;
            const int inv = kind == 1 || kind == 3 || kind == 4 || kind == 6;
            const fr_t* pre = r == 0 ? (kind == 2 ? tabs[i]->coset[0] : kind == 5 ? tabs[i]->coset_rev[0] : nullptr) : nullptr;
            const fr_t* post = r + 1 == m && kind == 6 ? tabs[i]->coset_rev[1] : nullptr;
            NttPassDesc d{(fr_t*)jobs[i].x, tabs[i]->tw[inv], pre, post, ks[i], pl.x, pl.y, pl.z, kind == 5, tiles};
            pds.push_back(d);
            tiles += (uint64_t)1 << (ks[i] - pl.y - pl.z);
        }
        round_tiles[r] = tiles;
    }
    round_begin[rounds] = pds.size();
    // the permutation of kinds 0..3: element-wise below 2^8 points (begin in
    // elements), else in 16 x 16 blocks (a second table, begin in blocks)
    std::vector<NttElemDesc> eds, bds;
    uint64_t elems = 0, blocks = 0;
    for (size_t i = 0; i < count; ++i) {
        const int kind = jobs[i].kind;
        if (kind > 3) continue;
        const NttElemDesc d{(fr_t*)jobs[i].x, (kind & 1) ? tabs[i]->ninv : nullptr, kind == 3 ? tabs[i]->coset[1] : nullptr,
                            ks[i], ks[i] >= 8 ? blocks : elems};
        if (ks[i] >= 8) {
            bds.push_back(d);
            blocks += (uint64_t)1 << (ks[i] - 8);
        } else {
            eds.push_back(d);
            elems += jobs[i].deg;
        }
    }
    // descriptors to the device
    size_t pbytes = pds.size() * sizeof(NttPassDesc), ebytes = eds.size() * sizeof(NttElemDesc),
           bbytes = bds.size() * sizeof(NttElemDesc);
    size_t eoff = (pbytes + 255) & ~size_t(255), boff = (eoff + ebytes + 255) & ~size_t(255);
    std::vector<char> host(boff + bbytes);
    std::memcpy(host.data(), pds.data(), pbytes);
    std::memcpy(host.data() + eoff, eds.data(), ebytes);
    std::memcpy(host.data() + boff, bds.data(), bbytes);
    const void* dev = nullptr;
    int rc = upload_desc(ctx, host.data(), host.size(), &dev);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const NttPassDesc* dpd = (const NttPassDesc*)dev;
    const NttElemDesc* ded = (const NttElemDesc*)((const char*)dev + eoff);
    const NttElemDesc* dbd = (const NttElemDesc*)((const char*)dev + boff);

    auto stages = [&]() {
        for (size_t i = 0; i < count; ++i) {   // 2- and 4-point transforms
            if (ks[i] == 0 || ks[i] >= 3) continue;
            const int kind = jobs[i].kind, inv = kind & 1;
            if (kind == 2)
                hipLaunchKernelGGL(k_mul_table, dim3(1), dim3(256), 0, st, (fr_t*)jobs[i].x, (const fr_t*)tabs[i]->coset[0],
                                   jobs[i].deg, (const fr_t*)nullptr);
            hipLaunchKernelGGL(k_ntt_tiny, dim3(1), dim3(64), 0, st, (fr_t*)jobs[i].x, (const fr_t*)tabs[i]->tw[inv], ks[i]);
        }
        for (size_t r = 0; r < rounds; ++r) {
            uint32_t np = (uint32_t)(round_begin[r + 1] - round_begin[r]);
            const dim3 grid((unsigned)round_tiles[r]);
            const NttPassDesc* rp = dpd + round_begin[r];
            if (LE == 1) {
                if (TB == 9) hipLaunchKernelGGL((k_ntt_pass_batch<9, 1>), grid, dim3(ntt_threads<9, 1>()), 0, st, rp, np);
                else if (TB == 10) hipLaunchKernelGGL((k_ntt_pass_batch<10, 1>), grid, dim3(ntt_threads<10, 1>()), 0, st, rp, np);
                else hipLaunchKernelGGL((k_ntt_pass_batch<11, 1>), grid, dim3(ntt_threads<11, 1>()), 0, st, rp, np);
            } else {
                if (TB == 9) hipLaunchKernelGGL((k_ntt_pass_batch<9, 3>), grid, dim3(ntt_threads<9, 3>()), 0, st, rp, np);
                else if (TB == 10) hipLaunchKernelGGL((k_ntt_pass_batch<10, 3>), grid, dim3(ntt_threads<10, 3>()), 0, st, rp, np);
                else hipLaunchKernelGGL((k_ntt_pass_batch<11, 3>), grid, dim3(ntt_threads<11, 3>()), 0, st, rp, np);
            }
        }
        if (elems)
            hipLaunchKernelGGL(k_bitrev_batch, dim3(blocks_for(elems, 256)), dim3(256), 0, st, ded, (uint32_t)eds.size(), elems);
        if (blocks)
            hipLaunchKernelGGL(k_bitrev_tiles_batch, dim3((unsigned)blocks), dim3(256), 0, st, dbd, (uint32_t)bds.size());
    };
    if (capture) {
        std::string key = "ntt:" + std::to_string(TB) + "/" + std::to_string(LE) + ":";
        for (size_t i = 0; i < count; ++i) {
            key += std::to_string(ks[i]) + "/" + std::to_string(jobs[i].kind) + ";";
            if (ks[i] > 0 && ks[i] < 3) key += std::to_string((uintptr_t)jobs[i].x) + ";";   // launched with direct arguments
        }
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
