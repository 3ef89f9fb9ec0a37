// Fq in nine 29-bit limbs for the G1 bucket accumulation (the MSM's hot loop).
//
// With 32-bit limbs every 32x32 partial product of a Montgomery product is a
// v_mad_u64_u32 plus a v_addc_co_u32 that folds its carry-out: two VALU
// instructions per partial product. With 29-bit limbs a 64-bit column
// accumulator cannot overflow (a column holds at most 9 products of two
// limbs below 2^30, 9 reduction products below 2^58 and the previous
// column's carry: < 2^63.5), so each partial product is one v_mad_u64_u32
// into the accumulator and nothing else: 162 multiply-adds per product
// instead of 128 multiply-adds and 128 carry folds.
//
// Montgomery form with R = 2^261 (nine full 29-bit reduction digits). Values
// are lazily reduced: a product returns a value below 2p whenever the two
// inputs' bounds, in units of p, multiply to less than 2^261 / p > 168 (the
// caller's invariant, checked in the point formulas' comments); sums and
// differences carry explicit multiples of p. The bases and the bucket sums
// stay in the library's R = 2^256 form in memory: from_r256 / to_r256
// convert with one product each.
#pragma once
#include "curve.hpp"
#include "field.hpp"

namespace hzk {

// acc += a * b as one v_mad_u64_u32 whose addend is the running column sum.
// As plain C++ the compiler re-associates each column into partial chains
// joined by extra 64-bit adds (17 per product); pinned this way those adds
// disappear, but the compiler also puts an s_nop between consecutive inline
// asm blocks (mad_row groups the interleaved products' multiplies of one
// column step into one block to have fewer of them), which only waves that
// share their SIMD with other waves hide: the accumulations (two or three
// waves per SIMD) use it, the latency-bound stages do not.
HZK_DEV void mad_pinned(uint64_t& acc, uint32_t a, uint32_t b) {
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc) : "v"(a), "v"(b) : "vcc");
}
template <bool PIN>
HZK_DEV void mad29(uint64_t& acc, uint32_t a, uint32_t b) {
    if constexpr (PIN) mad_pinned(acc, a, b);
    else acc += (uint64_t)a * b;
}
// acc[c] += x[c] * y[c] for the N interleaved products of one column step,
// pinned as one asm block (independent multiplies, no s_nop between them)
template <int N, bool PIN>
HZK_DEV void mad_row(uint64_t* acc, const uint32_t* x, const uint32_t* y) {
    if constexpr (!PIN || N > 3) {
#pragma unroll
        for (int c = 0; c < N; ++c) mad29<PIN>(acc[c], x[c], y[c]);
    } else if constexpr (N == 1) {
        asm("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc[0]) : "v"(x[0]), "v"(y[0]) : "vcc");
    } else if constexpr (N == 2) {
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\tv_mad_u64_u32 %1, vcc, %4, %5, %1"
            : "+v"(acc[0]), "+v"(acc[1]) : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]) : "vcc");
    } else {
        asm("v_mad_u64_u32 %0, vcc, %3, %4, %0\n\tv_mad_u64_u32 %1, vcc, %5, %6, %1\n\tv_mad_u64_u32 %2, vcc, %7, %8, %2"
            : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]) : "v"(x[0]), "v"(y[0]), "v"(x[1]), "v"(y[1]), "v"(x[2]), "v"(y[2]) : "vcc");
    }
}

struct fq29 {
    static constexpr uint32_t MASK = (1u << 29) - 1;
    static constexpr uint32_t P[9] = {0x187cfd47, 0x010460b6, 0x1c72a34f, 0x02d522d0, 0x1585d978,
                                      0x02db40c0, 0x00a6e141, 0x0e5c2634, 0x0030644e};
    static constexpr uint32_t M0 = 0x04866389;     // -p^-1 mod 2^29
    static constexpr uint32_t PINV = 0x1b799c77;   // p^-1 mod 2^29
    static constexpr uint32_t ONE[9] = {0x157ccc21, 0x141c2758, 0x185230d3, 0x014c0419, 0x0aa36fb9,
                                        0x1d4240ce, 0x11d54c07, 0x052ac7a8, 0x000dc836};   // 2^261 mod p
    static constexpr uint32_t TO[9] = {0x13349ca1, 0x1a5d84a8, 0x0a3e5cac, 0x100249e0, 0x12b951e8,
                                       0x0e92d304, 0x14cb95b3, 0x041b9d3d, 0x00058003};    // 2^266 mod p
    static constexpr uint32_t FROM[9] = {0x058f0d9d, 0x1aea1c6e, 0x11c2cf74, 0x11d651eb, 0x1462c0a7,
                                         0x11b7bc3c, 0x1cbd99ba, 0x183340fb, 0x000e0a77};  // 2^256 mod p
    // k p for the biases of subtractions
    static constexpr uint32_t P1[9] = {0x187cfd47, 0x010460b6, 0x1c72a34f, 0x02d522d0, 0x1585d978,
                                       0x02db40c0, 0x00a6e141, 0x0e5c2634, 0x0030644e};
    static constexpr uint32_t P2[9] = {0x10f9fa8e, 0x0208c16d, 0x18e5469e, 0x05aa45a1, 0x0b0bb2f0,
                                       0x05b68181, 0x014dc282, 0x1cb84c68, 0x0060c89c};
    static constexpr uint32_t P4[9] = {0x01f3f51c, 0x041182db, 0x11ca8d3c, 0x0b548b43, 0x161765e0,
                                       0x0b6d0302, 0x029b8504, 0x197098d0, 0x00c19139};
    static constexpr uint32_t P6[9] = {0x12edefaa, 0x061a4448, 0x0aafd3da, 0x10fed0e5, 0x012318d0,
                                       0x11238484, 0x03e94786, 0x1628e538, 0x012259d6};
    static constexpr uint32_t P18[9] = {0x18c9cefe, 0x124eccd9, 0x000f7b8e, 0x12fc72b0, 0x03694a71,
                                        0x136a8d8c, 0x0bbbd693, 0x027aafa8, 0x03670d84};
    static constexpr uint32_t P8[9] = {0x03e7ea38, 0x082305b6, 0x03951a78, 0x16a91687, 0x0c2ecbc0,
                                       0x16da0605, 0x05370a08, 0x12e131a0, 0x01832273};

    uint32_t l[9];

    static HZK_DEV fq29 zero() { fq29 r; for (int i = 0; i < 9; ++i) r.l[i] = 0; return r; }
    static HZK_DEV fq29 one() { fq29 r; for (int i = 0; i < 9; ++i) r.l[i] = ONE[i]; return r; }
    static HZK_DEV fq29 konst(const uint32_t* c) { fq29 r; for (int i = 0; i < 9; ++i) r.l[i] = c[i]; return r; }

    // repack 8 x 32-bit limbs (a value < 2^256) into 9 x 29-bit limbs
    static HZK_DEV fq29 unpack(const uint32_t* w) {
        fq29 r;
        r.l[0] = w[0] & MASK;
#pragma unroll
        for (int i = 1; i < 8; ++i) {
            const int b = 29 * i, q = b >> 5, s = b & 31;   // limb i starts at bit s of word q
            r.l[i] = (uint32_t)(((((uint64_t)w[q + 1]) << 32) | w[q]) >> s) & MASK;
        }
        r.l[8] = w[7] >> 8;   // bits 232..255
        return r;
    }
    // the same from two 16-byte words (low, high) with no array in between:
    // a small private array indexed in a loop can be placed in LDS by the
    // compiler, which turns every base load into LDS traffic
    static HZK_DEV fq29 unpack(const uint4& lo, const uint4& hi) {
        fq29 r;
        r.l[0] = lo.x & MASK;
        r.l[1] = ((lo.x >> 29) | (lo.y << 3)) & MASK;
        r.l[2] = ((lo.y >> 26) | (lo.z << 6)) & MASK;
        r.l[3] = ((lo.z >> 23) | (lo.w << 9)) & MASK;
        r.l[4] = ((lo.w >> 20) | (hi.x << 12)) & MASK;
        r.l[5] = ((hi.x >> 17) | (hi.y << 15)) & MASK;
        r.l[6] = ((hi.y >> 14) | (hi.z << 18)) & MASK;
        r.l[7] = ((hi.z >> 11) | (hi.w << 21)) & MASK;
        r.l[8] = hi.w >> 8;
        return r;
    }
    // normalized limbs of a value < 2^256 back to 8 x 32-bit limbs
    HZK_DEV void pack(uint32_t* w) const {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int b = 32 * q, i = b / 29, s = b % 29;   // word q starts at bit s of limb i
            uint64_t v = (uint64_t)l[i] >> s;
            v |= (uint64_t)l[i + 1] << (29 - s);
            if (i + 2 < 9 && 58 - s < 32) v |= (uint64_t)l[i + 2] << (58 - s);
            w[q] = (uint32_t)v;
        }
    }

    // a + b limb by limb, no carries: limbs below 2^30 for normalized inputs
    friend HZK_DEV fq29 operator+(const fq29& a, const fq29& b) {
        fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = a.l[i] + b.l[i];
        return r;
    }
    // a - b + K (K = kp, kp >= the value of b), normalized; limbs of a and b
    // below 2^31, values such that the result is below 2^261
    static HZK_DEV fq29 sub(const fq29& a, const fq29& b, const uint32_t* K) {
        fq29 r;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int32_t v = (int32_t)(a.l[i] + K[i] - b.l[i]) + c;
            r.l[i] = (uint32_t)v & MASK;
            c = v >> 29;
        }
        return r;
    }
    // kp - a for a value below kp, normalized
    static HZK_DEV fq29 neg(const fq29& a, const uint32_t* K) { return sub(zero(), a, K); }
    // carries propagated: normalized limbs of the same value (limbs below 2^31)
    HZK_DEV fq29 norm() const {
        fq29 r;
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const uint32_t v = l[i] + c;
            r.l[i] = v & MASK;
            c = v >> 29;
        }
        return r;
    }

    // a b 2^-261 mod p, lazily reduced (see the header); inputs' limbs below 2^30
    static HZK_DEV void mont_mul(fq29& r, const fq29& a, const fq29& b) {
        uint32_t m[9];
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
#pragma unroll
            for (int j = 0; j < k; ++j) acc += (uint64_t)m[j] * P[k - j];
#pragma unroll
            for (int j = 0; j <= k; ++j) acc += (uint64_t)a.l[j] * b.l[k - j];
            m[k] = ((uint32_t)acc * M0) & MASK;
            acc += (uint64_t)m[k] * P[0];
            acc >>= 29;
        }
#pragma unroll
        for (int k = 9; k < 17; ++k) {
#pragma unroll
            for (int j = k - 8; j < 9; ++j) acc += (uint64_t)a.l[j] * b.l[k - j];
#pragma unroll
            for (int j = k - 8; j < 9; ++j) acc += (uint64_t)m[j] * P[k - j];
            r.l[k - 9] = (uint32_t)acc & MASK;
            acc >>= 29;
        }
        r.l[8] = (uint32_t)acc;
    }
    // N independent products in one interleaved scan: each multiply-add of
    // product c depends only on product c's accumulator, so interleaving
    // hides the multiply-add latency that a single product's chain exposes
    template <int N, bool PIN = false>
    static HZK_DEV void mont_mul_n(fq29* r, const fq29* a, const fq29* b) {
        uint32_t m[N][9];
        uint64_t acc[N];
#pragma unroll
        for (int c = 0; c < N; ++c) acc[c] = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
#pragma unroll
            for (int j = 0; j < k; ++j)
                {
                    uint32_t x[N], y[N];
#pragma unroll
                    for (int c = 0; c < N; ++c) x[c] = m[c][j], y[c] = P[k - j];
                    mad_row<N, PIN>(acc, x, y);
                }
#pragma unroll
            for (int j = 0; j <= k; ++j)
                {
                    uint32_t x[N], y[N];
#pragma unroll
                    for (int c = 0; c < N; ++c) x[c] = a[c].l[j], y[c] = b[c].l[k - j];
                    mad_row<N, PIN>(acc, x, y);
                }
#pragma unroll
            for (int c = 0; c < N; ++c) {
                m[c][k] = ((uint32_t)acc[c] * M0) & MASK;
                mad29<PIN>(acc[c], m[c][k], P[0]);
                acc[c] >>= 29;
            }
        }
#pragma unroll
        for (int k = 9; k < 17; ++k) {
#pragma unroll
            for (int j = k - 8; j < 9; ++j)
                {
                    uint32_t x[N], y[N];
#pragma unroll
                    for (int c = 0; c < N; ++c) x[c] = a[c].l[j], y[c] = b[c].l[k - j];
                    mad_row<N, PIN>(acc, x, y);
                }
#pragma unroll
            for (int j = k - 8; j < 9; ++j)
                {
                    uint32_t x[N], y[N];
#pragma unroll
                    for (int c = 0; c < N; ++c) x[c] = m[c][j], y[c] = P[k - j];
                    mad_row<N, PIN>(acc, x, y);
                }
#pragma unroll
            for (int c = 0; c < N; ++c) {
                r[c].l[k - 9] = (uint32_t)acc[c] & MASK;
                acc[c] >>= 29;
            }
        }
#pragma unroll
        for (int c = 0; c < N; ++c) r[c].l[8] = (uint32_t)acc[c];
    }
    // N independent squarings, interleaved as in mont_mul_n: the column sum
    // sum_j a_j a_(k-j) taken as 2 sum_(j < k-j) a_j a_(k-j) + a_(k/2)^2 (the
    // same integer, so the same result) with the doubled limbs precomputed:
    // 45 limb products instead of 81. Inputs normalized (limbs below 2^29,
    // doubled below 2^30; a column stays below 2^62.2).
    template <int N, bool PIN = false>
    static HZK_DEV void mont_sqr_n(fq29* r, const fq29* a) {
        uint32_t m[N][9], d[N][9];
        uint64_t acc[N];
#pragma unroll
        for (int c = 0; c < N; ++c) {
            acc[c] = 0;
#pragma unroll
            for (int i = 0; i < 9; ++i) d[c][i] = a[c].l[i] << 1;
        }
#pragma unroll
        for (int k = 0; k < 17; ++k) {
            if (k < 9) {
#pragma unroll
                for (int j = 0; j < k; ++j)
#pragma unroll
                    for (int c = 0; c < N; ++c) mad29<PIN>(acc[c], m[c][j], P[k - j]);
            }
#pragma unroll
            for (int j = (k > 8 ? k - 8 : 0); 2 * j < k; ++j)
                {
                    uint32_t x[N], y[N];
#pragma unroll
                    for (int c = 0; c < N; ++c) x[c] = d[c][j], y[c] = a[c].l[k - j];
                    mad_row<N, PIN>(acc, x, y);
                }
            if ((k & 1) == 0)
#pragma unroll
                for (int c = 0; c < N; ++c) mad29<PIN>(acc[c], a[c].l[k / 2], a[c].l[k / 2]);
            if (k < 9) {
#pragma unroll
                for (int c = 0; c < N; ++c) {
                    m[c][k] = ((uint32_t)acc[c] * M0) & MASK;
                    mad29<PIN>(acc[c], m[c][k], P[0]);
                    acc[c] >>= 29;
                }
            } else {
#pragma unroll
                for (int j = k - 8; j < 9; ++j)
#pragma unroll
                    for (int c = 0; c < N; ++c) mad29<PIN>(acc[c], m[c][j], P[k - j]);
#pragma unroll
                for (int c = 0; c < N; ++c) {
                    r[c].l[k - 9] = (uint32_t)acc[c] & MASK;
                    acc[c] >>= 29;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < N; ++c) r[c].l[8] = (uint32_t)acc[c];
    }
    template <bool PIN = false>
    static HZK_DEV void sqr2(fq29& r0, const fq29& a0, fq29& r1, const fq29& a1) {
        const fq29 a[2] = {a0, a1};
        fq29 r[2];
        mont_sqr_n<2, PIN>(r, a);
        r0 = r[0], r1 = r[1];
    }
    static HZK_DEV fq29 sqr1(const fq29& a0) {
        fq29 r;
        mont_sqr_n<1>(&r, &a0);
        return r;
    }
    template <bool PIN = false>
    static HZK_DEV void mul2(fq29& r0, const fq29& a0, const fq29& b0, fq29& r1, const fq29& a1, const fq29& b1) {
        const fq29 a[2] = {a0, a1}, b[2] = {b0, b1};
        fq29 r[2];
        mont_mul_n<2, PIN>(r, a, b);
        r0 = r[0], r1 = r[1];
    }
    template <bool PIN = false>
    static HZK_DEV void mul3(fq29& r0, const fq29& a0, const fq29& b0, fq29& r1, const fq29& a1, const fq29& b1, fq29& r2,
                             const fq29& a2, const fq29& b2) {
        const fq29 a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2};
        fq29 r[3];
        mont_mul_n<3, PIN>(r, a, b);
        r0 = r[0], r1 = r[1], r2 = r[2];
    }
    friend HZK_DEV fq29 operator*(const fq29& a, const fq29& b) { fq29 r; mont_mul(r, a, b); return r; }
    HZK_DEV fq29 sqr() const { return *this * *this; }

    // a value below 2^261 reduced into [0, p): subtract q p with q from the
    // top limb (an underestimate by at most one), then p once more if needed
    HZK_DEV fq29 canonical() const {
        // q = floor(l8 / (P[8] + 1)) <= floor(value / p); the remainder is below 2p + ...
        fq29 r = *this;
        for (int round = 0; round < 2; ++round) {
            const uint32_t q = r.l[8] / (P[8] + 1);
            if (q) {
                int64_t c = 0;
                for (int i = 0; i < 9; ++i) {
                    const int64_t v = (int64_t)r.l[i] - (int64_t)q * P[i] + c;
                    r.l[i] = (uint32_t)v & MASK;
                    c = v >> 29;
                }
            }
        }
        // now below 2p (l8 <= P8 means value < (P8 + 1) 2^232 <= 2p): one conditional subtraction
        fq29 d;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int32_t v = (int32_t)r.l[i] - (int32_t)P[i] + c;
            d.l[i] = (uint32_t)v & MASK;
            c = v >> 29;
        }
        return c < 0 ? r : d;
    }

    // a normalized value below 2^261 brought below 2p: subtract q p with
    // q = floor(l8 / (P8 + 1)) <= floor(value / p); what remains is below
    // (P8 + 1 + q) 2^232 < 2p
    HZK_DEV fq29 reduce() const {
        const uint32_t q = l[8] / (P[8] + 1);
        fq29 r;
        int64_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int64_t v = (int64_t)l[i] - (int64_t)q * P[i] + c;
            r.l[i] = (uint32_t)v & MASK;
            c = v >> 29;
        }
        return r;
    }

    // value == 0 mod p, for a normalized value below 16p: v = k p exactly
    // when k = v p^-1 mod 2^29 is small and k p equals v
    HZK_DEV bool is_zero_mod_p() const {
        const uint32_t k = (l[0] * PINV) & MASK;
        if (k > 16) return false;
        if (k == 0) {
            uint32_t t = 0;
            for (int i = 0; i < 9; ++i) t |= l[i];
            return t == 0;
        }
        uint64_t c = 0;
        for (int i = 0; i < 9; ++i) {
            c += (uint64_t)k * P[i];
            if (((uint32_t)c & MASK) != l[i]) return false;
            c >>= 29;
        }
        return c == 0;
    }

    // R = 2^256 form (8 x 32-bit limbs, < 2^256) <-> this form: one product each
    static HZK_DEV fq29 from_r256(const fq_t& x) { return unpack(x.l) * konst(TO); }
    HZK_DEV fq_t to_r256() const {
        const fq29 v = (*this * konst(FROM)).canonical();
        fq_t r;
        v.pack(r.l);
        return r;
    }
};

// G1 bucket accumulator over fq29 in XYZZ coordinates (madd-2008-s), with the
// point at infinity as a flag. Invariants (values in units of p): X < 8,
// Y < 4, ZZ < 2, ZZZ < 2; affine inputs are below p.
struct g1acc29 {
    fq29 X, Y, ZZ, ZZZ;
    bool inf;

    HZK_DEV void set_inf() { inf = true; }
    HZK_DEV void set_affine(const fq29& x, const fq29& y) {
        X = x, Y = y, ZZ = fq29::one(), ZZZ = fq29::one();
        inf = false;
    }
    // this += (x, y) for a finite affine point (x, y < p), every special case
    // handled. When (x, y) equals this point (P = R = 0) the same code doubles
    // it: dbl-2008-s-1 on the XYZZ state is madd-2008-s with P = 2Y, R = 3X^2
    // and no PPP term in X3, so the rare case costs one squaring and a select
    // instead of a second formula's registers.
    HZK_DEV void add_affine(const fq29& x, const fq29& y) {
        if (inf) { set_affine(x, y); return; }
        fq29 U2, S2;
        fq29::mul2<true>(U2, x, ZZ, S2, y, ZZZ);               // 1 x 2 each
        fq29 P = fq29::sub(U2, X, fq29::P8);             // < 10
        fq29 R = fq29::sub(S2, Y, fq29::P4);             // < 6
        bool dbl = false;
        if (P.is_zero_mod_p()) {
            if (!R.is_zero_mod_p()) { set_inf(); return; }
            P = (Y + Y).norm();                          // 2Y < 8
            const fq29 xx = fq29::sqr1(X);                     // 8 x 8
            R = (xx + xx + xx).norm();                   // 3X^2 < 6
            dbl = true;
        }
        fq29 PP, RR;
        fq29::sqr2<true>(PP, P, RR, R);                  // 10 x 10, 6 x 6
        fq29 PPP, Q, ZZ3;
        fq29::mul3<true>(PPP, P, PP, Q, X, PP, ZZ3, ZZ, PP);   // 10 x 2, 8 x 2, 2 x 2
        fq29 T = Q + Q;
        if (!dbl) T = T + PPP;
        X = fq29::sub(RR, T, fq29::P6);                  // RR - (PPP + 2Q) + 6p < 8
        const fq29 D = fq29::sub(Q, X, fq29::P8);        // < 10
        fq29 t, YP, ZZZ3;
        fq29::mul3<true>(t, R, D, YP, Y, PPP, ZZZ3, ZZZ, PPP); // 6 x 10, 4 x 2, 2 x 2
        Y = fq29::sub(t, YP, fq29::P2);                  // < 4
        ZZ = ZZ3, ZZZ = ZZZ3;
    }
    // the R = 2^256 XYZZ point (canonical coordinates; infinity = all zero)
    HZK_DEV xyzz_t<fq_t> to_xyzz() const {
        if (inf) return xyzz_t<fq_t>::inf();
        return {X.to_r256(), Y.to_r256(), ZZ.to_r256(), ZZZ.to_r256()};
    }
    // The raw state, 36 words (X, Y, ZZ, ZZZ limbs; infinity has ZZ = 0, which
    // a finite point's ZZ = Z^2 never is): a plain store, so a lane that
    // finishes a bucket in the middle of its chunk holds the others up for a
    // few instructions instead of four conversions. The later stages (bucket
    // fix-up and reduction) read and write this form too; only the window
    // sums leave it (to_xyzz).
    static constexpr int RAW_WORDS = 36;
    HZK_DEV void to_words(uint32_t* w) const {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            w[i] = X.l[i], w[9 + i] = Y.l[i], w[18 + i] = inf ? 0u : ZZ.l[i], w[27 + i] = ZZZ.l[i];
        }
    }
    static HZK_DEV g1acc29 from_words(const uint32_t* w) {
        g1acc29 a;
        uint32_t zz = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            a.X.l[i] = w[i], a.Y.l[i] = w[9 + i], a.ZZ.l[i] = w[18 + i], a.ZZZ.l[i] = w[27 + i];
            zz |= w[18 + i];
        }
        a.inf = zz == 0;
        return a;
    }
    HZK_DEV void store_raw(uint4* dst) const {
        uint32_t w[RAW_WORDS];
        to_words(w);
#pragma unroll
        for (int q = 0; q < RAW_WORDS / 4; ++q) dst[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    }
    static HZK_DEV g1acc29 load_raw(const uint4* src) {
        uint32_t w[RAW_WORDS];
#pragma unroll
        for (int q = 0; q < RAW_WORDS / 4; ++q) {
            const uint4 v = src[q];
            w[4 * q] = v.x, w[4 * q + 1] = v.y, w[4 * q + 2] = v.z, w[4 * q + 3] = v.w;
        }
        return from_words(w);
    }

    // this += q for two states under the invariants (add-2008-s); when q
    // equals this point the same tail doubles it (dbl-2008-s-1 as in
    // add_affine: U1 -> X, S1 -> Y, the ZZ, ZZZ products -> this point's)
    HZK_DEV void add(const g1acc29& q) {
        if (q.inf) return;
        if (inf) { *this = q; return; }
        fq29 U1, U2, ZZ12, S1, S2, ZZZ12;
        fq29::mul3(U1, X, q.ZZ, U2, q.X, ZZ, ZZ12, ZZ, q.ZZ);        // 8 x 2, 8 x 2, 2 x 2
        fq29::mul3(S1, Y, q.ZZZ, S2, q.Y, ZZZ, ZZZ12, ZZZ, q.ZZZ);   // 4 x 2, 4 x 2, 2 x 2
        fq29 P = fq29::sub(U2, U1, fq29::P2);                        // < 4
        fq29 R = fq29::sub(S2, S1, fq29::P2);                        // < 4
        bool d = false;
        if (P.is_zero_mod_p()) {
            if (!R.is_zero_mod_p()) { set_inf(); return; }
            P = (Y + Y).norm();                                      // 2Y < 8
            const fq29 xx = fq29::sqr1(X);                                 // 8 x 8
            R = (xx + xx + xx).norm();                               // 3X^2 < 6
            U1 = X, S1 = Y, ZZ12 = ZZ, ZZZ12 = ZZZ;
            d = true;
        }
        finish(P, R, U1, S1, ZZ12, ZZZ12, d);
    }
    HZK_DEV void dbl() {
        if (inf) return;
        const fq29 xx = fq29::sqr1(X);
        finish((Y + Y).norm(), (xx + xx + xx).norm(), X, Y, ZZ, ZZZ, true);
    }
    // the common tail: P < 8, R < 6, U1 < 8, S1 < 4, ZZ12, ZZZ12 < 2
    HZK_DEV void finish(const fq29& P, const fq29& R, const fq29& U1, const fq29& S1, const fq29& ZZ12, const fq29& ZZZ12,
                        bool d) {
        fq29 PP, RR;
        fq29::sqr2(PP, P, RR, R);                              // 8 x 8, 6 x 6
        fq29 PPP, Q, ZZn;
        fq29::mul3(PPP, P, PP, Q, U1, PP, ZZn, ZZ12, PP);            // 8 x 2, 8 x 2, 2 x 2
        fq29 T = Q + Q;
        if (!d) T = T + PPP;
        X = fq29::sub(RR, T, fq29::P6);                              // < 8
        const fq29 D = fq29::sub(Q, X, fq29::P8);                    // < 10
        fq29 t, YP, ZZZn;
        fq29::mul3(t, R, D, YP, S1, PPP, ZZZn, ZZZ12, PPP);          // 6 x 10, 4 x 2, 2 x 2
        Y = fq29::sub(t, YP, fq29::P2);                              // < 4
        ZZ = ZZn, ZZZ = ZZZn;
    }
};

// Fq2 = Fq[u]/(u^2 + 1) over fq29 components
struct fq2_29 {
    fq29 c0, c1;
    static HZK_DEV fq2_29 one() { return {fq29::one(), fq29::zero()}; }
    friend HZK_DEV fq2_29 operator+(const fq2_29& a, const fq2_29& b) { return {a.c0 + b.c0, a.c1 + b.c1}; }
    static HZK_DEV fq2_29 sub(const fq2_29& a, const fq2_29& b, const uint32_t* K) {
        return {fq29::sub(a.c0, b.c0, K), fq29::sub(a.c1, b.c1, K)};
    }
    HZK_DEV fq2_29 reduce() const { return {c0.reduce(), c1.reduce()}; }
    HZK_DEV fq2_29 norm() const { return {c0.norm(), c1.norm()}; }
    HZK_DEV bool is_zero_mod_p() const { return c0.is_zero_mod_p() && c1.is_zero_mod_p(); }
    // Karatsuba, 3 interleaved products: c0 = a0 b0 - a1 b1 (< 4p), c1 =
    // (a0 + a1)(b0 + b1) - a0 b0 - a1 b1 (< 6p); inputs normalized, component
    // bounds with (A0 + A1)(B0 + B1) < 168
    friend HZK_DEV fq2_29 operator*(const fq2_29& a, const fq2_29& b) {
        const fq29 v0 = a.c0 * b.c0, v1 = a.c1 * b.c1, v2 = (a.c0 + a.c1) * (b.c0 + b.c1);
        return {fq29::sub(v0, v1, fq29::P2), fq29::sub(v2, v0 + v1, fq29::P4)};
    }
    // the same products in one interleaved scan (three independent chains of
    // multiply-adds): for kernels at one wave per SIMD, where a single
    // product's dependent chain would leave the SIMD waiting
    template <bool PIN = false>
    static HZK_DEV fq2_29 mul_x(const fq2_29& a, const fq2_29& b) {
        fq29 v0, v1, v2;
        fq29::mul3<PIN>(v0, a.c0, b.c0, v1, a.c1, b.c1, v2, a.c0 + a.c1, b.c0 + b.c1);
        return {fq29::sub(v0, v1, fq29::P2), fq29::sub(v2, v0 + v1, fq29::P4)};
    }
    template <bool PIN = false>
    HZK_DEV fq2_29 sqr_x() const {
        fq29 t, ab;
        fq29::mul2<PIN>(t, c0 + c1, fq29::sub(c0, c1, fq29::P2), ab, c0, c1);
        return {t, (ab + ab).norm()};
    }
    // (a0 + a1)(a0 - a1) and 2 a0 a1, 2 interleaved products; components below 2p:
    // c0 < 2p, c1 < 4p
    HZK_DEV fq2_29 sqr() const {
        const fq29 t = (c0 + c1) * fq29::sub(c0, c1, fq29::P2), ab = c0 * c1;
        return {t, (ab + ab).norm()};
    }
};

// G2 bucket accumulator over fq2_29 (madd-2008-s), infinity as a flag.
// Invariants (component bounds in units of p): X, Y < 2 (reduced after every
// subtraction), ZZ, ZZZ < 6 (products); affine inputs below p. Every
// Karatsuba product then sees (A0 + A1)(B0 + B1) <= 100.
struct g2acc29 {
    fq2_29 X, Y, ZZ, ZZZ;
    bool inf;

    HZK_DEV void set_inf() { inf = true; }
    HZK_DEV void set_affine(const fq2_29& x, const fq2_29& y) {
        X = x, Y = y, ZZ = fq2_29::one(), ZZZ = fq2_29::one();
        inf = false;
    }
    // this += (x, y), every special case handled; a doubling (x, y equal to
    // this point) runs the same code with P = 2Y, R = 3X^2 and no PPP term in
    // X3 (dbl-2008-s-1); ordered for short live ranges (the lane state is 72
    // registers)
    HZK_DEV void add_affine(const fq2_29& x, const fq2_29& y) {
        if (inf) { set_affine(x, y); return; }
        fq2_29 P = fq2_29::sub(fq2_29::mul_x<true>(x, ZZ), X, fq29::P2).reduce();                // U2 = x ZZ: (2)(10); P < 2
        fq2_29 R = fq2_29::sub(fq2_29::mul_x<true>(y, ZZZ), Y, fq29::P2).reduce();               // S2 = y ZZZ; R < 2
        bool dbl = false;
        if (P.is_zero_mod_p()) {
            if (!R.is_zero_mod_p()) { set_inf(); return; }
            P = (Y + Y).norm().reduce();                                     // 2Y < 2
            const fq2_29 xx = X.template sqr_x<true>();
            R = (xx + xx + xx).norm().reduce();                              // 3X^2 < 2
            dbl = true;
        }
        const fq2_29 PP = P.template sqr_x<true>();                                           // < 4
        const fq2_29 PPP = fq2_29::mul_x<true>(P, PP);                                           // (4)(6), < 6
        ZZ = fq2_29::mul_x<true>(ZZ, PP);                                                        // (10)(6)
        const fq2_29 Q = fq2_29::mul_x<true>(X, PP);                                             // (4)(6), < 6
        fq2_29 T = Q + Q;
        if (!dbl) T = T + PPP;
        X = fq2_29::sub(R.template sqr_x<true>(), T, fq29::P18).reduce();                     // RR - (PPP + 2Q) + 18p, then < 2
        const fq2_29 t = fq2_29::mul_x<true>(R, fq2_29::sub(Q, X, fq29::P2).reduce());           // R (Q - X3): (4)(4)
        Y = fq2_29::sub(t, fq2_29::mul_x<true>(Y, PPP), fq29::P6).reduce();                      // Y PPP: (4)(10); Y3 < 2
        ZZZ = fq2_29::mul_x<true>(ZZZ, PPP);                                                     // (10)(10)
    }
    // raw state: 72 words, infinity has ZZ = 0 (read back by load_raw)
    static constexpr int RAW_WORDS = 72;
    HZK_DEV void to_words(uint32_t* w) const {
        const fq29* c[8] = {&X.c0, &X.c1, &Y.c0, &Y.c1, &ZZ.c0, &ZZ.c1, &ZZZ.c0, &ZZZ.c1};
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int i = 0; i < 9; ++i) w[9 * k + i] = (inf && (k == 4 || k == 5)) ? 0u : c[k]->l[i];
    }
    static HZK_DEV g2acc29 from_words(const uint32_t* w) {
        g2acc29 a;
        fq29* c[8] = {&a.X.c0, &a.X.c1, &a.Y.c0, &a.Y.c1, &a.ZZ.c0, &a.ZZ.c1, &a.ZZZ.c0, &a.ZZZ.c1};
        uint32_t zz = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int i = 0; i < 9; ++i) c[k]->l[i] = w[9 * k + i];
#pragma unroll
        for (int i = 0; i < 18; ++i) zz |= w[36 + i];
        a.inf = zz == 0;
        return a;
    }
    HZK_DEV void store_raw(uint4* dst) const {
        uint32_t w[RAW_WORDS];
        to_words(w);
#pragma unroll
        for (int q = 0; q < RAW_WORDS / 4; ++q) dst[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    }
    static HZK_DEV g2acc29 load_raw(const uint4* src) {
        uint32_t w[RAW_WORDS];
#pragma unroll
        for (int q = 0; q < RAW_WORDS / 4; ++q) {
            const uint4 v = src[q];
            w[4 * q] = v.x, w[4 * q + 1] = v.y, w[4 * q + 2] = v.z, w[4 * q + 3] = v.w;
        }
        return from_words(w);
    }
    // the R = 2^256 XYZZ point (canonical components; infinity = all zero)
    HZK_DEV xyzz_t<fq2_t> to_xyzz() const {
        if (inf) return xyzz_t<fq2_t>::inf();
        return {{X.c0.to_r256(), X.c1.to_r256()}, {Y.c0.to_r256(), Y.c1.to_r256()},
                {ZZ.c0.to_r256(), ZZ.c1.to_r256()}, {ZZZ.c0.to_r256(), ZZZ.c1.to_r256()}};
    }

    // this += q for two states under the invariants (add-2008-s); when q
    // equals this point the same tail doubles it (dbl-2008-s-1 with U1 -> X,
    // S1 -> Y and the ZZ, ZZZ products -> this point's). Interleaved
    // products (mul_x): the callers run at one wave per SIMD.
    HZK_DEV void add(const g2acc29& q) {
        if (q.inf) return;
        if (inf) { *this = q; return; }
        fq2_29 U1 = fq2_29::mul_x(X, q.ZZ);                                  // (4)(10) -> c0 < 4, c1 < 6
        fq2_29 P = fq2_29::sub(fq2_29::mul_x(q.X, ZZ), U1, fq29::P6).reduce();   // U2 - U1, < 2
        fq2_29 S1 = fq2_29::mul_x(Y, q.ZZZ);
        fq2_29 R = fq2_29::sub(fq2_29::mul_x(q.Y, ZZZ), S1, fq29::P6).reduce();  // S2 - S1, < 2
        bool d = false;
        if (P.is_zero_mod_p()) {
            if (!R.is_zero_mod_p()) { set_inf(); return; }
            P = (Y + Y).norm().reduce();
            const fq2_29 xx = X.sqr_x();
            R = (xx + xx + xx).norm().reduce();
            U1 = X, S1 = Y;
            d = true;
        }
        const fq2_29 ZZ12 = d ? ZZ : fq2_29::mul_x(ZZ, q.ZZ);               // (10)(10) -> (4, 6)
        const fq2_29 ZZZ12 = d ? ZZZ : fq2_29::mul_x(ZZZ, q.ZZZ);
        finish(P, R, U1, S1, ZZ12, ZZZ12, d);
    }
    HZK_DEV void dbl() {
        if (inf) return;
        const fq2_29 xx = X.sqr_x();
        finish((Y + Y).norm().reduce(), (xx + xx + xx).norm().reduce(), X, Y, ZZ, ZZZ, true);
    }
    // the common tail: P, R < 2; U1, S1, ZZ12, ZZZ12 with c0 < 4, c1 < 6
    HZK_DEV void finish(const fq2_29& P, const fq2_29& R, const fq2_29& U1, const fq2_29& S1, const fq2_29& ZZ12,
                        const fq2_29& ZZZ12, bool d) {
        const fq2_29 PP = P.sqr_x();                                         // (2, 4)
        const fq2_29 PPP = fq2_29::mul_x(P, PP);                             // (4)(6)
        ZZ = fq2_29::mul_x(ZZ12, PP);                                        // (10)(6)
        ZZZ = fq2_29::mul_x(ZZZ12, PPP);                                     // (10)(10)
        const fq2_29 Q = fq2_29::mul_x(U1, PP);                              // (10)(6)
        fq2_29 T = Q + Q;
        if (!d) T = T + PPP;                                                 // c0 < 12, c1 < 18
        X = fq2_29::sub(R.sqr_x(), T, fq29::P18).reduce();                   // < 2
        const fq2_29 t = fq2_29::mul_x(R, fq2_29::sub(Q, X, fq29::P2).reduce());   // (4)(4)
        Y = fq2_29::sub(t, fq2_29::mul_x(S1, PPP), fq29::P6).reduce();       // S1 PPP: (10)(10); Y3 < 2
    }
};

}  // namespace hzk
