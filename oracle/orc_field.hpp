// ORACLE — test infrastructure only. CPU restatement of the reference's host
// arithmetic (libff subset vendored under include/host-arith). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this code,
// and only as the checker; the product never links it.
//
// Prime fields Fq, Fr of alt_bn128 in Montgomery form, R = 2^256, 4x64-bit
// little-endian limbs, every value fully reduced into [0, p). This is the byte
// layout of libff's Fp_model::mont_repr (include/host-arith/field/fp.cuh:42,
// bigint<4>::data at include/host-arith/field/bigint.cuh:29) and of the device
// mont_t (include/device-arith/field/mont_t.cuh:33), so buffers are shared
// verbatim between the two.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>

namespace orc {

typedef unsigned __int128 u128;

struct FieldParams {
    uint64_t p[4];      // modulus
    uint64_t inv;       // -p^{-1} mod 2^64
    uint64_t r2[4];     // R^2 mod p
    uint64_t one[4];    // R mod p
};

// Constants restate src/host-arith/curve/alt_bn128/alt_bn128_init.cu:40-61
// (moduli, Rsquared, inv) and include/device-arith/curve/alt_bn128_t.cuh:11-45.
inline constexpr FieldParams FQ_PARAMS = {
    {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0x87d20782e4866389ULL,
    {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL},
    {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL},
};
inline constexpr FieldParams FR_PARAMS = {
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0xc2e1f593efffffffULL,
    {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL},
    {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL},
};

static inline int cmp4(const uint64_t a[4], const uint64_t b[4]) {
    for (int i = 3; i >= 0; --i) {
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    }
    return 0;
}
static inline uint64_t add4(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
    u128 c = 0;
    for (int i = 0; i < 4; ++i) { c += (u128)a[i] + b[i]; r[i] = (uint64_t)c; c >>= 64; }
    return (uint64_t)c;
}
static inline uint64_t sub4(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a[i] - b[i] - borrow;
        r[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) ? 1 : 0;
    }
    return borrow;
}

template <const FieldParams& P>
struct Fp {
    uint64_t v[4];

    static Fp zero() { Fp r; std::memset(r.v, 0, 32); return r; }
    static Fp one() { Fp r; std::memcpy(r.v, P.one, 32); return r; }
    static Fp from_raw(const uint64_t m[4]) { Fp r; std::memcpy(r.v, m, 32); return r; }
    static const uint64_t* modulus() { return P.p; }

    bool is_zero() const { return (v[0] | v[1] | v[2] | v[3]) == 0; }
    bool operator==(const Fp& o) const { return std::memcmp(v, o.v, 32) == 0; }
    bool operator!=(const Fp& o) const { return !(*this == o); }

    // Montgomery product a*b*2^-256 mod p: CIOS (Koc, Acar, Kaliski 1996), the
    // algorithm libff's mul_reduce implements (include/host-arith/field/fp.cuh:199-244).
    static void mont_mul(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
        uint64_t t[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            u128 c = 0;
            for (int j = 0; j < 4; ++j) {
                c = (u128)a[j] * b[i] + t[j] + (uint64_t)(c >> 64);
                t[j] = (uint64_t)c;
            }
            c = (u128)t[4] + (uint64_t)(c >> 64);
            t[4] = (uint64_t)c;
            t[5] = (uint64_t)(c >> 64);
            uint64_t m = t[0] * P.inv;
            c = (u128)m * P.p[0] + t[0];
            for (int j = 1; j < 4; ++j) {
                c = (u128)m * P.p[j] + t[j] + (uint64_t)(c >> 64);
                t[j - 1] = (uint64_t)c;
            }
            c = (u128)t[4] + (uint64_t)(c >> 64);
            t[3] = (uint64_t)c;
            t[4] = t[5] + (uint64_t)(c >> 64);
        }
        uint64_t s[4];
        uint64_t borrow = sub4(s, t, P.p);
        if (t[4] || !borrow) std::memcpy(r, s, 32);
        else std::memcpy(r, t, 32);
    }

    Fp operator*(const Fp& o) const { Fp r; mont_mul(r.v, v, o.v); return r; }
    Fp& operator*=(const Fp& o) { mont_mul(v, v, o.v); return *this; }
    Fp squared() const { return *this * *this; }

    Fp operator+(const Fp& o) const {
        Fp r; uint64_t c = add4(r.v, v, o.v);
        uint64_t s[4];
        uint64_t b = sub4(s, r.v, P.p);
        if (c || !b) std::memcpy(r.v, s, 32);
        return r;
    }
    Fp operator-(const Fp& o) const {
        Fp r; uint64_t b = sub4(r.v, v, o.v);
        if (b) add4(r.v, r.v, P.p);
        return r;
    }
    Fp operator-() const { return Fp::zero() - *this; }
    Fp& operator+=(const Fp& o) { *this = *this + o; return *this; }
    Fp& operator-=(const Fp& o) { *this = *this - o; return *this; }
    Fp dbl() const { return *this + *this; }

    // Canonical (non-Montgomery) integer: multiply by 1 (fp.cuh:371-382 as_bigint).
    void to_canonical(uint64_t out[4]) const {
        const uint64_t one_raw[4] = {1, 0, 0, 0};
        mont_mul(out, v, one_raw);
    }
    // Montgomery form of a canonical integer x < p: x*R^2*R^-1 (fp.cuh:333-341).
    static Fp from_canonical(const uint64_t x[4]) {
        Fp r; mont_mul(r.v, x, P.r2); return r;
    }
    static Fp from_u64(uint64_t x) {
        uint64_t t[4] = {x, 0, 0, 0};
        return from_canonical(t);
    }

    // this^e for a little-endian exponent of nlimbs 64-bit words (exp.cuh:43-70).
    Fp pow(const uint64_t* e, int nlimbs) const {
        Fp acc = one();
        bool found = false;
        for (int i = nlimbs * 64 - 1; i >= 0; --i) {
            if (found) acc = acc.squared();
            if ((e[i / 64] >> (i % 64)) & 1) { found = true; acc = acc * *this; }
        }
        return acc;
    }
    // Fermat inverse a^(p-2) (fp.cuh:766-780); 0 maps to 0.
    Fp inverse() const {
        uint64_t e[4]; const uint64_t two[4] = {2, 0, 0, 0};
        sub4(e, P.p, two);
        return pow(e, 4);
    }

    static Fp from_hex(const char* hex);   // canonical big-endian hex -> Montgomery
    std::string to_hex() const;            // canonical big-endian hex, 64 digits
};

template <const FieldParams& P>
Fp<P> Fp<P>::from_hex(const char* hex) {
    uint64_t x[4] = {0, 0, 0, 0};
    size_t len = std::strlen(hex);
    for (size_t i = 0; i < len; ++i) {
        char ch = hex[len - 1 - i];
        uint64_t d = (ch >= '0' && ch <= '9') ? ch - '0' : (ch >= 'a' && ch <= 'f') ? ch - 'a' + 10 : ch - 'A' + 10;
        x[i / 16] |= d << (4 * (i % 16));
    }
    return from_canonical(x);
}

template <const FieldParams& P>
std::string Fp<P>::to_hex() const {
    uint64_t c[4]; to_canonical(c);
    static const char* digits = "0123456789abcdef";
    std::string s(64, '0');
    for (int i = 0; i < 64; ++i) s[63 - i] = digits[(c[i / 16] >> (4 * (i % 16))) & 15];
    return s;
}

typedef Fp<FQ_PARAMS> Fq;
typedef Fp<FR_PARAMS> Fr;

// Fq2 = Fq[u]/(u^2 + 1): the non-residue is p-1 (alt_bn128_init.cu:77).
// Layout c0 | c1 (include/host-arith/field/fp2.cuh:41).
struct Fq2 {
    Fq c0, c1;
    static Fq2 zero() { return {Fq::zero(), Fq::zero()}; }
    static Fq2 one() { return {Fq::one(), Fq::zero()}; }
    bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
    bool operator==(const Fq2& o) const { return c0 == o.c0 && c1 == o.c1; }
    bool operator!=(const Fq2& o) const { return !(*this == o); }
    Fq2 operator+(const Fq2& o) const { return {c0 + o.c0, c1 + o.c1}; }
    Fq2 operator-(const Fq2& o) const { return {c0 - o.c0, c1 - o.c1}; }
    Fq2 operator-() const { return {-c0, -c1}; }
    Fq2 dbl() const { return {c0.dbl(), c1.dbl()}; }
    // Karatsuba (fp2.cuh:171-183): (a0+a1u)(b0+b1u) = a0b0 - a1b1 + ((a0+a1)(b0+b1) - a0b0 - a1b1)u
    Fq2 operator*(const Fq2& o) const {
        Fq aa = c0 * o.c0, bb = c1 * o.c1;
        return {aa - bb, (c0 + c1) * (o.c0 + o.c1) - aa - bb};
    }
    Fq2 operator*(const Fq& s) const { return {c0 * s, c1 * s}; }
    Fq2& operator*=(const Fq2& o) { *this = *this * o; return *this; }
    Fq2& operator+=(const Fq2& o) { *this = *this + o; return *this; }
    Fq2& operator-=(const Fq2& o) { *this = *this - o; return *this; }
    // complex squaring (fp2.cuh:222-233 squared_complex)
    Fq2 squared() const {
        Fq ab = c0 * c1;
        return {(c0 + c1) * (c0 - c1), ab.dbl()};
    }
    // inverse via the norm (fp2.cuh:235-249): 1/(a+bu) = (a - bu)/(a^2 + b^2)
    Fq2 inverse() const {
        Fq t = (c0.squared() + c1.squared()).inverse();
        return {c0 * t, -(c1 * t)};
    }
    Fq2 conj() const { return {c0, -c1}; }
    Fq2 mul_by_xi() const {  // * (9 + u)
        Fq nine = Fq::from_u64(9);
        return {c0 * nine - c1, c1 * nine + c0};
    }
};

}  // namespace orc
