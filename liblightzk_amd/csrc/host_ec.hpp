// Host-side alt_bn128 arithmetic for the sequential tails of the GPU path:
// the MSM window combine (c doublings per window, Horner), affine
// normalisation of results, and the Groth16 phase-6 assembly, which the
// reference also performs on the host (include/protocol/groth16.cuh:210-215).
// These are a few thousand dependent field operations per call: one GPU wave
// needs ~1 us per 256-bit Montgomery product, a CPU core ~20 ns, so they run
// here. Values use the same 32-byte Montgomery layout as the device.
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

namespace hzk {
namespace host {

typedef unsigned __int128 u128;

struct Params { uint64_t p[4]; uint64_t inv; uint64_t one[4]; uint64_t r2[4]; };
inline constexpr Params FQ = {
    {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0x87d20782e4866389ULL,
    {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL},
    {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL}};
inline constexpr Params FR = {
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0xc2e1f593efffffffULL,
    {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL},
    {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL}};

template <const Params& P>
struct Fp {
    uint64_t v[4];
    static Fp zero() { return Fp{{0, 0, 0, 0}}; }
    static Fp one() { Fp r; std::memcpy(r.v, P.one, 32); return r; }
    static Fp load(const void* src) { Fp r; std::memcpy(r.v, src, 32); return r; }
    void store(void* dst) const { std::memcpy(dst, v, 32); }
    bool is_zero() const { return (v[0] | v[1] | v[2] | v[3]) == 0; }
    bool operator==(const Fp& o) const { return std::memcmp(v, o.v, 32) == 0; }
    bool operator!=(const Fp& o) const { return !(*this == o); }

    static bool geq_p(const uint64_t* t) {
        for (int i = 3; i >= 0; --i) { if (t[i] != P.p[i]) return t[i] > P.p[i]; }
        return true;
    }
    static void sub_p(uint64_t* t) {
        uint64_t bw = 0;
        for (int i = 0; i < 4; ++i) { u128 d = (u128)t[i] - P.p[i] - bw; t[i] = (uint64_t)d; bw = (uint64_t)(d >> 127); }
    }
    // Montgomery product, operand scanning with interleaved reduction
    friend Fp operator*(const Fp& a, const Fp& b) {
        uint64_t t[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            u128 c = 0;
            for (int j = 0; j < 4; ++j) { c = (u128)a.v[j] * b.v[i] + t[j] + (uint64_t)(c >> 64); t[j] = (uint64_t)c; }
            uint64_t top = (uint64_t)(c >> 64);
            uint64_t m = t[0] * P.inv;
            c = (u128)m * P.p[0] + t[0];
            for (int j = 1; j < 4; ++j) { c = (u128)m * P.p[j] + t[j] + (uint64_t)(c >> 64); t[j - 1] = (uint64_t)c; }
            c = (u128)t[4] + top + (uint64_t)(c >> 64);
            t[3] = (uint64_t)c;
            t[4] = (uint64_t)(c >> 64);
        }
        if (t[4] || geq_p(t)) sub_p(t);
        Fp r; std::memcpy(r.v, t, 32); return r;
    }
    friend Fp operator+(const Fp& a, const Fp& b) {
        Fp r; u128 c = 0;
        for (int i = 0; i < 4; ++i) { c += (u128)a.v[i] + b.v[i]; r.v[i] = (uint64_t)c; c >>= 64; }
        if (geq_p(r.v)) sub_p(r.v);
        return r;
    }
    friend Fp operator-(const Fp& a, const Fp& b) {
        Fp r; uint64_t bw = 0;
        for (int i = 0; i < 4; ++i) { u128 d = (u128)a.v[i] - b.v[i] - bw; r.v[i] = (uint64_t)d; bw = (uint64_t)(d >> 127); }
        if (bw) { u128 c = 0; for (int i = 0; i < 4; ++i) { c += (u128)r.v[i] + P.p[i]; r.v[i] = (uint64_t)c; c >>= 64; } }
        return r;
    }
    Fp neg() const { return zero() - *this; }
    Fp dbl() const { return *this + *this; }
    Fp sqr() const { return *this * *this; }
    Fp pow(const uint64_t e[4]) const {
        Fp acc = one();
        for (int i = 255; i >= 0; --i) { acc = acc.sqr(); if ((e[i / 64] >> (i % 64)) & 1) acc = acc * *this; }
        return acc;
    }
    Fp inverse() const {   // a^(p-2); 0 -> 0
        uint64_t e[4] = {P.p[0] - 2, P.p[1], P.p[2], P.p[3]};
        return pow(e);
    }
    void canonical(uint64_t out[4]) const {   // leave Montgomery form
        Fp one_raw{{1, 0, 0, 0}};
        Fp r = *this * one_raw;
        std::memcpy(out, r.v, 32);
    }
};
typedef Fp<FQ> Fq;
typedef Fp<FR> Fr;

struct Fq2 {
    Fq c0, c1;
    static Fq2 zero() { return {Fq::zero(), Fq::zero()}; }
    static Fq2 one() { return {Fq::one(), Fq::zero()}; }
    static Fq2 load(const void* s) { const char* p = (const char*)s; return {Fq::load(p), Fq::load(p + 32)}; }
    void store(void* d) const { char* p = (char*)d; c0.store(p); c1.store(p + 32); }
    bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
    bool operator==(const Fq2& o) const { return c0 == o.c0 && c1 == o.c1; }
    friend Fq2 operator+(const Fq2& a, const Fq2& b) { return {a.c0 + b.c0, a.c1 + b.c1}; }
    friend Fq2 operator-(const Fq2& a, const Fq2& b) { return {a.c0 - b.c0, a.c1 - b.c1}; }
    friend Fq2 operator*(const Fq2& a, const Fq2& b) {
        Fq aa = a.c0 * b.c0, bb = a.c1 * b.c1;
        return {aa - bb, (a.c0 + a.c1) * (b.c0 + b.c1) - aa - bb};
    }
    Fq2 neg() const { return {c0.neg(), c1.neg()}; }
    Fq2 dbl() const { return {c0.dbl(), c1.dbl()}; }
    Fq2 sqr() const { Fq ab = c0 * c1; return {(c0 + c1) * (c0 - c1), ab.dbl()}; }
    Fq2 inverse() const { Fq t = (c0.sqr() + c1.sqr()).inverse(); return {c0 * t, (c1 * t).neg()}; }
    static constexpr int BYTES = 64;
};

// Field-generic XYZZ point (x = X/ZZ, y = Y/ZZZ; infinity ZZ = 0), the same
// coordinates the device buckets use, so device partial sums load verbatim.
template <class F>
struct Xyzz {
    F X, Y, ZZ, ZZZ;
    static Xyzz inf() { return {F::zero(), F::zero(), F::zero(), F::zero()}; }
    bool is_inf() const { return ZZ.is_zero(); }
    static Xyzz from_affine(const F& x, const F& y) {
        if (x.is_zero() && y.is_zero()) return inf();
        return {x, y, F::one(), F::one()};
    }
    Xyzz dbl() const {
        if (is_inf()) return *this;
        F U = Y.dbl(), V = U.sqr(), W = U * V, S = X * V, xx = X.sqr();
        F M = xx.dbl() + xx;
        F X3 = M.sqr() - S.dbl();
        F Y3 = M * (S - X3) - W * Y;
        return {X3, Y3, V * ZZ, W * ZZZ};
    }
    Xyzz add(const Xyzz& o) const {
        if (o.is_inf()) return *this;
        if (is_inf()) return o;
        F U1 = X * o.ZZ, U2 = o.X * ZZ, S1 = Y * o.ZZZ, S2 = o.Y * ZZZ;
        F P = U2 - U1, R = S2 - S1;
        if (P.is_zero()) return R.is_zero() ? dbl() : inf();
        F PP = P.sqr(), PPP = P * PP, Q = U1 * PP;
        F X3 = R.sqr() - PPP - Q.dbl();
        F Y3 = R * (Q - X3) - S1 * PPP;
        return {X3, Y3, ZZ * o.ZZ * PP, ZZZ * o.ZZZ * PPP};
    }
    Xyzz neg() const { return {X, Y.neg(), ZZ, ZZZ}; }
    // k * this for a canonical little-endian scalar
    Xyzz mul(const uint64_t k[4]) const {
        Xyzz acc = inf();
        for (int i = 255; i >= 0; --i) { acc = acc.dbl(); if ((k[i / 64] >> (i % 64)) & 1) acc = acc.add(*this); }
        return acc;
    }
    // affine (x, y); false for infinity
    bool to_affine(F& x, F& y) const {
        if (is_inf()) { x = F::zero(); y = F::zero(); return false; }
        F inv = (ZZ * ZZZ).inverse();      // 1/(ZZ ZZZ)
        x = X * (inv * ZZZ);                 // X / ZZ
        y = Y * (inv * ZZ);                  // Y / ZZZ
        return true;
    }
    // normalised Jacobian (x, y, 1) or libff's infinity (0, 1, 0)
    void store_jacobian(void* dst) const {
        char* d = (char*)dst;
        const int B = sizeof(F);
        F x, y;
        if (to_affine(x, y)) { x.store(d); y.store(d + B); F::one().store(d + 2 * B); }
        else { F::zero().store(d); F::one().store(d + B); F::zero().store(d + 2 * B); }
    }
    void store_affine(void* dst) const {
        char* d = (char*)dst;
        F x, y;
        to_affine(x, y);
        x.store(d); y.store(d + sizeof(F));
    }
    static Xyzz load_affine(const void* src) {
        const char* s = (const char*)src;
        return from_affine(F::load(s), F::load(s + sizeof(F)));
    }
    static Xyzz load_xyzz(const void* src) {
        const char* s = (const char*)src;
        const int B = sizeof(F);
        return {F::load(s), F::load(s + B), F::load(s + 2 * B), F::load(s + 3 * B)};
    }
};

// k * P for a point P fixed for many products (a proving key's delta and
// beta): 32 byte windows, entry (w, b) = b 2^(8w) P for b = 1..255, so a
// product is at most 32 additions and no doublings (double-and-add: 256
// doublings and ~128 additions)
template <class F>
struct FixedBase {
    std::vector<Xyzz<F>> t;   // 32 x 256 entries, b = 0 unused
    void build(const Xyzz<F>& p) {
        t.assign(32 * 256, Xyzz<F>::inf());
        Xyzz<F> base = p;
        for (int w = 0; w < 32; ++w) {
            Xyzz<F> acc = Xyzz<F>::inf();
            for (int b = 1; b < 256; ++b) {
                acc = acc.add(base);
                t[w * 256 + b] = acc;
            }
            for (int q = 0; q < 8; ++q) base = base.dbl();
        }
    }
    bool ready() const { return !t.empty(); }
    Xyzz<F> mul(const uint64_t k[4]) const {
        Xyzz<F> acc = Xyzz<F>::inf();
        for (int w = 0; w < 32; ++w) {
            const uint32_t b = (uint32_t)(k[w / 8] >> (8 * (w % 8))) & 0xffu;
            if (b) acc = acc.add(t[w * 256 + b]);
        }
        return acc;
    }
};

typedef Xyzz<Fq> G1x;
typedef Xyzz<Fq2> G2x;

// Generators (libff alt_bn128_init.cu:130-132, :191-195), Montgomery form.
inline G1x g1_generator() {
    Fq x = Fq::one();
    Fq y = x.dbl();
    return G1x::from_affine(x, y);
}
inline Fq fq_from_canonical(const uint64_t c[4]) { Fq t; std::memcpy(t.v, c, 32); Fq r2; std::memcpy(r2.v, FQ.r2, 32); return t * r2; }
inline G2x g2_generator() {
    static const uint64_t xc0[4] = {0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL, 0x1800deef121f1e76ULL};
    static const uint64_t xc1[4] = {0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL, 0x198e9393920d483aULL};
    static const uint64_t yc0[4] = {0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL, 0x12c85ea5db8c6debULL};
    static const uint64_t yc1[4] = {0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL, 0x090689d0585ff075ULL};
    Fq2 x{fq_from_canonical(xc0), fq_from_canonical(xc1)};
    Fq2 y{fq_from_canonical(yc0), fq_from_canonical(yc1)};
    return G2x::from_affine(x, y);
}

// Montgomery's trick: affine (x, y) of every point with one inversion.
template <class F>
void batch_affine(const std::vector<Xyzz<F>>& pts, std::vector<F>& xs, std::vector<F>& ys) {
    size_t n = pts.size();
    xs.assign(n, F::zero());
    ys.assign(n, F::zero());
    std::vector<F> pre(n);
    F acc = F::one();
    for (size_t i = 0; i < n; ++i) {
        pre[i] = acc;
        if (!pts[i].is_inf()) acc = acc * (pts[i].ZZ * pts[i].ZZZ);
    }
    F inv = acc.inverse();
    for (size_t i = n; i-- > 0;) {
        if (pts[i].is_inf()) continue;
        F d = pts[i].ZZ * pts[i].ZZZ;
        F di = inv * pre[i];   // 1 / (ZZ ZZZ)
        inv = inv * d;
        xs[i] = pts[i].X * (di * pts[i].ZZZ);
        ys[i] = pts[i].Y * (di * pts[i].ZZ);
    }
}

}  // namespace host
}  // namespace hzk
