// Device field arithmetic for alt_bn128 on gfx950 (CDNA4, wave64).
//
// Fq and Fr are 256-bit Montgomery fields (R = 2^256) held as 8 x 32-bit
// little-endian limbs, every value fully reduced into [0, p). The byte layout
// equals the reference's device mont_t (include/device-arith/field/mont_t.cuh:33)
// and libff's bigint<4> (include/host-arith/field/bigint.cuh:29), so host and
// device buffers are interchangeable.
//
// Multiplication is the finely integrated product-scanning Montgomery product
// (see fp::mont_mul): every 32x32 partial product is one v_mad_u64_u32 whose
// carry-out is folded by one v_addc_co_u32; adds and subtracts are
// v_add_co_u32/v_addc_co_u32 carry chains through VCC. No MFMA: nothing here
// is a floating-point contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hzk {

#define HZK_DEV __device__ __forceinline__

struct FqParams {
    // p = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
    static constexpr uint32_t MOD[8] = {0xd87cfd47, 0x3c208c16, 0x6871ca8d, 0x97816a91,
                                        0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72};
    static constexpr uint32_t M0 = 0xe4866389;   // -p^-1 mod 2^32
    static constexpr uint32_t ONE[8] = {0xc58f0d9d, 0xd35d438d, 0xf5c70b3d, 0x0a78eb28,
                                        0x7879462c, 0x666ea36f, 0x9a07df2f, 0x0e0a77c1};   // R mod p
    static constexpr uint32_t R2[8] = {0x538afa89, 0xf32cfc5b, 0xd44501fb, 0xb5e71911,
                                       0x0a417ff6, 0x47ab1eff, 0xcab8351f, 0x06d89f71};    // R^2 mod p
};
struct FrParams {
    // r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
    static constexpr uint32_t MOD[8] = {0xf0000001, 0x43e1f593, 0x79b97091, 0x2833e848,
                                        0x8181585d, 0xb85045b6, 0xe131a029, 0x30644e72};
    static constexpr uint32_t M0 = 0xefffffff;
    static constexpr uint32_t ONE[8] = {0x4ffffffb, 0xac96341c, 0x9f60cd29, 0x36fc7695,
                                        0x7879462e, 0x666ea36f, 0x9a07df2f, 0x0e0a77c1};
    static constexpr uint32_t R2[8] = {0xae216da7, 0x1bb8e645, 0xe35c59e3, 0x53fe3ab1,
                                       0x53bb8085, 0x8c49833d, 0x7f4e44a5, 0x0216d0b1};
};

// Carry-folding multiply-accumulate steps of 2 or 3 independent Montgomery
// products, interleaved: every chain's v_mad_u64_u32 first, then every
// chain's v_addc_co_u32. Each chain carries through its own SGPR pair, so the
// two wait states gfx950 needs between a VALU SGPR write and its use as
// carry-in are filled by the other chains' multiplies (one s_nop covers the
// two-chain case), and the compiler sees independent dependency chains
// instead of one chain serialised through VCC.
HZK_DEV void mac2_vv(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
    uint64_t c0, c1;
    asm("v_mad_u64_u32 %0, %4, %6, %7, %0\n\t"
        "v_mad_u64_u32 %1, %5, %8, %9, %1\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %2, %4, 0, %2, %4\n\t"
        "v_addc_co_u32_e64 %3, %5, 0, %3, %5"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(top[0]), "+v"(top[1]), "=&s"(c0), "=&s"(c1)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1));
}
HZK_DEV void mac2_vs(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t x1, uint32_t ys) {
    uint64_t c0, c1;
    asm("v_mad_u64_u32 %0, %4, %6, %8, %0\n\t"
        "v_mad_u64_u32 %1, %5, %7, %8, %1\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %2, %4, 0, %2, %4\n\t"
        "v_addc_co_u32_e64 %3, %5, 0, %3, %5"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(top[0]), "+v"(top[1]), "=&s"(c0), "=&s"(c1)
        : "v"(x0), "v"(x1), "s"(ys));
}
HZK_DEV void mac3_vv(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t x2,
                     uint32_t y2) {
    uint64_t c0, c1, c2;
    asm("v_mad_u64_u32 %0, %6, %9, %10, %0\n\t"
        "v_mad_u64_u32 %1, %7, %11, %12, %1\n\t"
        "v_mad_u64_u32 %2, %8, %13, %14, %2\n\t"
        "v_addc_co_u32_e64 %3, %6, 0, %3, %6\n\t"
        "v_addc_co_u32_e64 %4, %7, 0, %4, %7\n\t"
        "v_addc_co_u32_e64 %5, %8, 0, %5, %8"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(top[0]), "+v"(top[1]), "+v"(top[2]), "=&s"(c0), "=&s"(c1),
          "=&s"(c2)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1), "v"(x2), "v"(y2));
}
HZK_DEV void mac3_vs(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t ys) {
    uint64_t c0, c1, c2;
    asm("v_mad_u64_u32 %0, %6, %9, %12, %0\n\t"
        "v_mad_u64_u32 %1, %7, %10, %12, %1\n\t"
        "v_mad_u64_u32 %2, %8, %11, %12, %2\n\t"
        "v_addc_co_u32_e64 %3, %6, 0, %3, %6\n\t"
        "v_addc_co_u32_e64 %4, %7, 0, %4, %7\n\t"
        "v_addc_co_u32_e64 %5, %8, 0, %5, %8"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(top[0]), "+v"(top[1]), "+v"(top[2]), "=&s"(c0), "=&s"(c1),
          "=&s"(c2)
        : "v"(x0), "v"(x1), "v"(x2), "s"(ys));
}

// The same steps for the first product of a column: the carry word is
// written (top = carry-out) instead of accumulated, so the previous column's
// carry word never has to be cleared.
HZK_DEV void mac2_vv_first(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
    uint64_t c0, c1;
    asm("v_mad_u64_u32 %0, %4, %6, %7, %0\n\t"
        "v_mad_u64_u32 %1, %5, %8, %9, %1\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %2, %4, 0, 0, %4\n\t"
        "v_addc_co_u32_e64 %3, %5, 0, 0, %5"
        : "+v"(acc[0]), "+v"(acc[1]), "=&v"(top[0]), "=&v"(top[1]), "=&s"(c0), "=&s"(c1)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1));
}
HZK_DEV void mac2_vs_first(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t x1, uint32_t ys) {
    uint64_t c0, c1;
    asm("v_mad_u64_u32 %0, %4, %6, %8, %0\n\t"
        "v_mad_u64_u32 %1, %5, %7, %8, %1\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32_e64 %2, %4, 0, 0, %4\n\t"
        "v_addc_co_u32_e64 %3, %5, 0, 0, %5"
        : "+v"(acc[0]), "+v"(acc[1]), "=&v"(top[0]), "=&v"(top[1]), "=&s"(c0), "=&s"(c1)
        : "v"(x0), "v"(x1), "s"(ys));
}
HZK_DEV void mac3_vv_first(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t x2,
                           uint32_t y2) {
    uint64_t c0, c1, c2;
    asm("v_mad_u64_u32 %0, %6, %9, %10, %0\n\t"
        "v_mad_u64_u32 %1, %7, %11, %12, %1\n\t"
        "v_mad_u64_u32 %2, %8, %13, %14, %2\n\t"
        "v_addc_co_u32_e64 %3, %6, 0, 0, %6\n\t"
        "v_addc_co_u32_e64 %4, %7, 0, 0, %7\n\t"
        "v_addc_co_u32_e64 %5, %8, 0, 0, %8"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "=&v"(top[0]), "=&v"(top[1]), "=&v"(top[2]), "=&s"(c0), "=&s"(c1),
          "=&s"(c2)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1), "v"(x2), "v"(y2));
}
HZK_DEV void mac3_vs_first(uint64_t* acc, uint32_t* top, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t ys) {
    uint64_t c0, c1, c2;
    asm("v_mad_u64_u32 %0, %6, %9, %12, %0\n\t"
        "v_mad_u64_u32 %1, %7, %10, %12, %1\n\t"
        "v_mad_u64_u32 %2, %8, %11, %12, %2\n\t"
        "v_addc_co_u32_e64 %3, %6, 0, 0, %6\n\t"
        "v_addc_co_u32_e64 %4, %7, 0, 0, %7\n\t"
        "v_addc_co_u32_e64 %5, %8, 0, 0, %8"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "=&v"(top[0]), "=&v"(top[1]), "=&v"(top[2]), "=&s"(c0), "=&s"(c1),
          "=&s"(c2)
        : "v"(x0), "v"(x1), "v"(x2), "s"(ys));
}

// x * y mod 2^32 as the low half of a v_mad_u64_u32 (the reduction digits
// m_k); measured slightly faster in the MSM loops than v_mul_lo_u32
HZK_DEV uint32_t mul_lo_mad(uint32_t x, uint32_t ys) {
    uint64_t r, c;
    asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(c) : "v"(x), "s"(ys));
    return (uint32_t)r;
}

template <class P>
struct fp {
    uint32_t l[8];

    static HZK_DEV fp zero() { fp r; for (int i = 0; i < 8; ++i) r.l[i] = 0; return r; }
    static HZK_DEV fp one() { fp r; for (int i = 0; i < 8; ++i) r.l[i] = P::ONE[i]; return r; }
    static HZK_DEV fp raw(const uint32_t* s) { fp r; for (int i = 0; i < 8; ++i) r.l[i] = s[i]; return r; }

    HZK_DEV bool is_zero() const {
        uint32_t t = l[0];
        for (int i = 1; i < 8; ++i) t |= l[i];
        return t == 0;
    }
    HZK_DEV bool operator==(const fp& o) const {
        uint32_t t = 0;
        for (int i = 0; i < 8; ++i) t |= l[i] ^ o.l[i];
        return t == 0;
    }
    HZK_DEV bool operator!=(const fp& o) const { return !(*this == o); }

    // r = a - p if a >= p else a  (a < 2p)
    static HZK_DEV void final_sub(uint32_t* a) {
        uint32_t d[8];
        uint32_t borrow = 0;
        for (int i = 0; i < 8; ++i) d[i] = __builtin_subc(a[i], MOD_(i), borrow, &borrow);
        // borrow == 1 means a < p: keep a
        for (int i = 0; i < 8; ++i) a[i] = borrow ? a[i] : d[i];
    }
    static HZK_DEV constexpr uint32_t MOD_(int i) { return P::MOD[i]; }

    friend HZK_DEV fp operator+(const fp& a, const fp& b) {
        fp r;
        uint32_t c = 0;
        for (int i = 0; i < 8; ++i) r.l[i] = __builtin_addc(a.l[i], b.l[i], c, &c);
        final_sub(r.l);   // a + b < 2p < 2^255: no carry out of the top limb
        return r;
    }
    friend HZK_DEV fp operator-(const fp& a, const fp& b) {
        fp r;
        uint32_t bw = 0;
        for (int i = 0; i < 8; ++i) r.l[i] = __builtin_subc(a.l[i], b.l[i], bw, &bw);
        uint32_t mask = 0u - bw;
        uint32_t c = 0;
        for (int i = 0; i < 8; ++i) r.l[i] = __builtin_addc(r.l[i], P::MOD[i] & mask, c, &c);
        return r;
    }
    HZK_DEV fp dbl() const { return *this + *this; }
    HZK_DEV fp neg() const {
        // p - a, and 0 stays 0
        fp r;
        uint32_t bw = 0;
        for (int i = 0; i < 8; ++i) r.l[i] = __builtin_subc(P::MOD[i], l[i], bw, &bw);
        uint32_t nz = is_zero() ? 0u : 0xffffffffu;
        for (int i = 0; i < 8; ++i) r.l[i] &= nz;
        return r;
    }
    HZK_DEV fp cneg(bool flag) const { return flag ? neg() : *this; }

    // Montgomery product a*b*2^-256 mod p by finely integrated product
    // scanning (FIPS; Koc, Acar, Kaliski 1996): column k of the 16-column
    // product accumulates every a_j*b_(k-j) and m_j*p_(k-j) into a 64-bit
    // accumulator plus a carry word. Each step is one v_mad_u64_u32 whose
    // carry-out (VCC) is folded into the carry word by v_addc_co_u32, so the
    // accumulator never leaves its register pair. The low 8 columns produce
    // the reduction digits m_k = acc_lo * (-p^-1) mod 2^32, which clear the
    // column; the high 8 columns are the result (< 2p), then one conditional
    // subtraction makes it canonical.
    // gfx950 needs two wait states between a VALU write of VCC (the mad's
    // carry-out) and a VALU read of it as carry-in; the compiler inserts the
    // same s_nop between its own carry chains, but cannot see inside asm.
    static HZK_DEV void mac(uint64_t& acc, uint32_t& top, uint32_t x, uint32_t y) {
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
            "s_nop 1\n\t"
            "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
            : "+v"(acc), "+v"(top)
            : "v"(x), "v"(y)
            : "vcc");
    }
    static HZK_DEV void mac_s(uint64_t& acc, uint32_t& top, uint32_t x, uint32_t ys) {
        asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
            "s_nop 1\n\t"
            "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
            : "+v"(acc), "+v"(top)
            : "v"(x), "s"(ys)
            : "vcc");
    }
    static HZK_DEV void mont_mul(uint32_t* r, const uint32_t* a, const uint32_t* b) {
        uint32_t m[8], t[8];
        uint64_t acc = 0;
        uint32_t top = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
#pragma unroll
            for (int j = 0; j < k; ++j) mac_s(acc, top, m[j], P::MOD[k - j]);
#pragma unroll
            for (int j = 0; j <= k; ++j) mac(acc, top, a[j], b[k - j]);
            m[k] = (uint32_t)acc * P::M0;
            mac_s(acc, top, m[k], P::MOD[0]);
            acc = (acc >> 32) | ((uint64_t)top << 32);
            top = 0;
        }
#pragma unroll
        for (int k = 8; k < 15; ++k) {
#pragma unroll
            for (int j = k - 7; j < 8; ++j) mac(acc, top, a[j], b[k - j]);
#pragma unroll
            for (int j = k - 7; j < 8; ++j) mac_s(acc, top, m[j], P::MOD[k - j]);
            t[k - 8] = (uint32_t)acc;
            acc = (acc >> 32) | ((uint64_t)top << 32);
            top = 0;
        }
        t[7] = (uint32_t)acc;   // result < 2p < 2^255: nothing above column 15
        final_sub(t);
        for (int i = 0; i < 8; ++i) r[i] = t[i];
    }

    // N = 2 or 3 independent Montgomery products in one interleaved scan
    // (same FIPS schedule as mont_mul, chain c computing r[c] = a[c] b[c] / R)
    template <int N>
    static HZK_DEV void mont_mul_n(uint32_t (*r)[8], const uint32_t* const* a, const uint32_t* const* b) {
        static_assert(N == 2 || N == 3, "2 or 3 chains");
        uint32_t m[N][8], t[N][8];
        uint64_t acc[N];
        uint32_t top[N];   // written by each column's first product
#pragma unroll
        for (int c = 0; c < N; ++c) acc[c] = 0;
        // first: the column's first product (its carry word starts from the carry-out)
        auto vv = [&](int i, int j, bool first) {
            if constexpr (N == 2) {
                if (first) mac2_vv_first(acc, top, a[0][i], b[0][j], a[1][i], b[1][j]);
                else mac2_vv(acc, top, a[0][i], b[0][j], a[1][i], b[1][j]);
            } else {
                if (first) mac3_vv_first(acc, top, a[0][i], b[0][j], a[1][i], b[1][j], a[2][i], b[2][j]);
                else mac3_vv(acc, top, a[0][i], b[0][j], a[1][i], b[1][j], a[2][i], b[2][j]);
            }
        };
        auto vs = [&](int j, uint32_t ys, bool first) {
            if constexpr (N == 2) {
                if (first) mac2_vs_first(acc, top, m[0][j], m[1][j], ys);
                else mac2_vs(acc, top, m[0][j], m[1][j], ys);
            } else {
                if (first) mac3_vs_first(acc, top, m[0][j], m[1][j], m[2][j], ys);
                else mac3_vs(acc, top, m[0][j], m[1][j], m[2][j], ys);
            }
        };
        auto shift = [&]() {
#pragma unroll
            for (int c = 0; c < N; ++c) acc[c] = (acc[c] >> 32) | ((uint64_t)top[c] << 32);
        };
#pragma unroll
        for (int k = 0; k < 8; ++k) {
#pragma unroll
            for (int j = 0; j < k; ++j) vs(j, P::MOD[k - j], j == 0);
#pragma unroll
            for (int j = 0; j <= k; ++j) vv(j, k - j, k == 0 && j == 0);
#pragma unroll
            for (int c = 0; c < N; ++c) m[c][k] = mul_lo_mad((uint32_t)acc[c], P::M0);
            vs(k, P::MOD[0], false);
            shift();
        }
#pragma unroll
        for (int k = 8; k < 15; ++k) {
#pragma unroll
            for (int j = k - 7; j < 8; ++j) vv(j, k - j, j == k - 7);
#pragma unroll
            for (int j = k - 7; j < 8; ++j) vs(j, P::MOD[k - j], false);
#pragma unroll
            for (int c = 0; c < N; ++c) t[c][k - 8] = (uint32_t)acc[c];
            shift();
        }
#pragma unroll
        for (int c = 0; c < N; ++c) {
            t[c][7] = (uint32_t)acc[c];
            final_sub(t[c]);
#pragma unroll
            for (int i = 0; i < 8; ++i) r[c][i] = t[c][i];
        }
    }
    // r0 = a0 b0, r1 = a1 b1 (and r2 = a2 b2) as interleaved chains
    static HZK_DEV void mul2(fp& r0, const fp& a0, const fp& b0, fp& r1, const fp& a1, const fp& b1) {
        const uint32_t* a[2] = {a0.l, a1.l};
        const uint32_t* b[2] = {b0.l, b1.l};
        uint32_t r[2][8];
        mont_mul_n<2>(r, a, b);
#pragma unroll
        for (int i = 0; i < 8; ++i) { r0.l[i] = r[0][i]; r1.l[i] = r[1][i]; }
    }
    static HZK_DEV void mul3(fp& r0, const fp& a0, const fp& b0, fp& r1, const fp& a1, const fp& b1, fp& r2, const fp& a2,
                             const fp& b2) {
        const uint32_t* a[3] = {a0.l, a1.l, a2.l};
        const uint32_t* b[3] = {b0.l, b1.l, b2.l};
        uint32_t r[3][8];
        mont_mul_n<3>(r, a, b);
#pragma unroll
        for (int i = 0; i < 8; ++i) { r0.l[i] = r[0][i]; r1.l[i] = r[1][i]; r2.l[i] = r[2][i]; }
    }

    friend HZK_DEV fp operator*(const fp& a, const fp& b) { fp r; mont_mul(r.l, a.l, b.l); return r; }
    HZK_DEV fp& operator*=(const fp& b) { mont_mul(l, l, b.l); return *this; }
    HZK_DEV fp& operator+=(const fp& b) { *this = *this + b; return *this; }
    HZK_DEV fp& operator-=(const fp& b) { *this = *this - b; return *this; }
    HZK_DEV fp sqr() const { fp r; mont_mul(r.l, l, l); return r; }
    static HZK_DEV void sqr2(fp& r0, const fp& a0, fp& r1, const fp& a1) { mul2(r0, a0, a0, r1, a1, a1); }

    // Montgomery form <-> canonical integer
    HZK_DEV fp from_mont() const {   // a * 1 * R^-1
        uint32_t one_raw[8] = {1, 0, 0, 0, 0, 0, 0, 0};
        fp r; mont_mul(r.l, l, one_raw); return r;
    }
    HZK_DEV fp to_mont() const { fp r; mont_mul(r.l, l, P::R2); return r; }

    // a^e for a little-endian exponent given as 8 x 32-bit words
    HZK_DEV fp pow(const uint32_t* e) const {
        fp acc = one();
        for (int i = 255; i >= 0; --i) {
            acc = acc.sqr();
            if ((e[i >> 5] >> (i & 31)) & 1) acc = acc * *this;
        }
        return acc;
    }
    // Fermat inverse a^(p-2); 0 maps to 0
    HZK_DEV fp inverse() const {
        uint32_t e[8];
        uint32_t bw = 0;
        e[0] = __builtin_subc(P::MOD[0], 2u, bw, &bw);
        for (int i = 1; i < 8; ++i) e[i] = __builtin_subc(P::MOD[i], 0u, bw, &bw);
        return pow(e);
    }
};

typedef fp<FqParams> fq_t;
typedef fp<FrParams> fr_t;

// Fq2 = Fq[u]/(u^2 + 1), one element per lane (no cross-lane traffic). The
// Fq2 product and square are inlined: as out-of-line calls they cost every
// caller a stack frame of saved registers (432 bytes of scratch per lane in
// the G2 accumulation) and measured 19 % slower than the larger inlined loop.
struct fq2_t {
    fq_t c0, c1;
    static HZK_DEV fq2_t zero() { return {fq_t::zero(), fq_t::zero()}; }
    static HZK_DEV fq2_t one() { return {fq_t::one(), fq_t::zero()}; }
    HZK_DEV bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
    HZK_DEV bool operator==(const fq2_t& o) const { return c0 == o.c0 && c1 == o.c1; }
    friend HZK_DEV fq2_t operator+(const fq2_t& a, const fq2_t& b) { return {a.c0 + b.c0, a.c1 + b.c1}; }
    friend HZK_DEV fq2_t operator-(const fq2_t& a, const fq2_t& b) { return {a.c0 - b.c0, a.c1 - b.c1}; }
    HZK_DEV fq2_t dbl() const { return {c0.dbl(), c1.dbl()}; }
    HZK_DEV fq2_t neg() const { return {c0.neg(), c1.neg()}; }
    HZK_DEV fq2_t cneg(bool f) const { return f ? neg() : *this; }
    // Karatsuba: 3 Fq multiplications, interleaved
    static HZK_DEV fq2_t mul_call(fq2_t a, fq2_t b) {
        fq_t aa, bb, s;
        fq_t::mul3(aa, a.c0, b.c0, bb, a.c1, b.c1, s, a.c0 + a.c1, b.c0 + b.c1);
        return {aa - bb, s - aa - bb};
    }
    // complex squaring: 2 Fq multiplications, interleaved
    static HZK_DEV fq2_t sqr_call(fq2_t a) {
        fq_t ab, t;
        fq_t::mul2(ab, a.c0, a.c1, t, a.c0 + a.c1, a.c0 - a.c1);
        return {t, ab.dbl()};
    }
    friend HZK_DEV fq2_t operator*(const fq2_t& a, const fq2_t& b) { return mul_call(a, b); }
    static HZK_DEV void mul2(fq2_t& r0, const fq2_t& a0, const fq2_t& b0, fq2_t& r1, const fq2_t& a1, const fq2_t& b1) {
        r0 = mul_call(a0, b0);
        r1 = mul_call(a1, b1);
    }
    static HZK_DEV void mul3(fq2_t& r0, const fq2_t& a0, const fq2_t& b0, fq2_t& r1, const fq2_t& a1, const fq2_t& b1,
                             fq2_t& r2, const fq2_t& a2, const fq2_t& b2) {
        r0 = mul_call(a0, b0);
        r1 = mul_call(a1, b1);
        r2 = mul_call(a2, b2);
    }
    static HZK_DEV void sqr2(fq2_t& r0, const fq2_t& a0, fq2_t& r1, const fq2_t& a1) {
        r0 = sqr_call(a0);
        r1 = sqr_call(a1);
    }
    HZK_DEV fq2_t& operator*=(const fq2_t& b) { *this = *this * b; return *this; }
    HZK_DEV fq2_t& operator+=(const fq2_t& b) { *this = *this + b; return *this; }
    HZK_DEV fq2_t& operator-=(const fq2_t& b) { *this = *this - b; return *this; }
    HZK_DEV fq2_t sqr() const { return sqr_call(*this); }
    HZK_DEV fq2_t inverse() const {
        fq_t t = (c0.sqr() + c1.sqr()).inverse();
        return {c0 * t, (c1 * t).neg()};
    }
};

}  // namespace hzk
