// Type names a caller of the reference's operators and protocol class passes
// as template arguments, as layout-only value types for host compilers.
//
// * alt_bn128::fp_t / fr_t / g1_t / g2_t stand for the reference's device
//   types (include/device-arith/curve/alt_bn128_t.cuh:53-80, :231-258): an
//   element is 32 bytes (8 x 32-bit limbs, little endian, Montgomery form with
//   R = 2^256), a G1 record is X | Y | Z (96 bytes), a G2 device record is one
//   lane's share X.ci | Y.ci | Z.ci (96 bytes; a point is two records, c0
//   then c1). Here they only name byte layouts: all arithmetic happens in the
//   library behind include/hzk.h.
// * libff::alt_bn128_Fr / _G1 / _G2 and init_alt_bn128_params stand for the
//   reference's host types (include/host-arith/field/fp.cuh,
//   include/host-arith/curve/alt_bn128/alt_bn128_g1.cuh, _g2.cuh) with their
//   byte layouts (bigint<4> limbs, bigint.cuh:29; G1 X | Y | Z 96 bytes, G2
//   X | Y | Z over Fq2 192 bytes) and the members the operator call sites use:
//   one(), zero(), random_element() (Fr) and print(). random_element() draws from
//   one seeded stream (xoshiro256** seeded by SplitMix64, rejection sampling
//   below the modulus after clearing bits 254 and 255, as fp.cuh:789-816
//   does with a random_device), so runs repeat; reseed with
//   libff::alt_bn128_seed().
// * ntt_config / alt_bn128_ntt_config stand for
//   include/host-arith/ntt-params.cuh:101-109. The library derives its roots
//   of unity itself, so the string-table pointers are null.
//
// A caller that links the reference's own host arithmetic defines
// LIGHTZK_EXTERNAL_LIBFF before including this header and brings its own
// libff types, ntt_config and alt_bn128_ntt_config; the operator shims need
// only their sizes (32, 96 and 192 bytes) and FieldT::one().
#pragma once
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

namespace alt_bn128 {
struct fp_t { uint32_t limb[8]; };
struct fr_t { uint32_t limb[8]; };
struct g1_t { fp_t X, Y, Z; };
struct g2_t { fp_t X, Y, Z; };   // one lane record of a G2 point (c0 or c1 parts)
static_assert(sizeof(fr_t) == 32 && sizeof(g1_t) == 96 && sizeof(g2_t) == 96, "reference device layouts");
}  // namespace alt_bn128

#ifndef LIGHTZK_EXTERNAL_LIBFF
namespace libff {

namespace detail {
// scalar field modulus r and R mod r (Montgomery one), little-endian 64-bit limbs
inline constexpr uint64_t FR_MOD[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                                       0x30644e72e131a029ULL};
inline constexpr uint64_t FR_ONE[4] = {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL,
                                       0x0e0a77c19a07df2fULL};
inline constexpr uint64_t FQ_ONE[4] = {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL,
                                       0x0e0a77c19a07df2fULL};
// generators in Montgomery form: G1 = (1, 2), G2 = (x0 + x1 u, y0 + y1 u)
// (the standard alt_bn128 generators, alt_bn128_init.cu:130-132, :191-195)
inline constexpr uint64_t G1_Y[4] = {0xa6ba871b8b1e1b3aULL, 0x14f1d651eb8e167bULL, 0xccdd46def0f28c58ULL,
                                     0x1c14ef83340fbe5eULL};
inline constexpr uint64_t G2_XY[4][4] = {
    {0x8e83b5d102bc2026ULL, 0xdceb1935497b0172ULL, 0xfbb8264797811adfULL, 0x19573841af96503bULL},
    {0xafb4737da84c6140ULL, 0x6043dd5a5802d8c4ULL, 0x09e950fc52a02f86ULL, 0x14fef0833aea7b6bULL},
    {0x619dfa9d886be9f6ULL, 0xfe7fd297f59e9b78ULL, 0xff9e1a62231b7dfeULL, 0x28fd7eebae9e4206ULL},
    {0x64095b56c71856eeULL, 0xdc57f922327d3cbbULL, 0x55f935be33351076ULL, 0x0da4a0e693fd6482ULL}};

struct Stream {
    uint64_t s[4];
    std::mutex mu;
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    void seed(uint64_t v) {
        for (auto& w : s) w = splitmix(v);
    }
    uint64_t next() {   // xoshiro256**
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    Stream() { seed(0x4c4947485a4bULL); }
};
inline Stream& stream() {
    static Stream st;
    return st;
}
inline bool below(const uint64_t* a, const uint64_t* m) {
    for (int i = 3; i >= 0; --i)
        if (a[i] != m[i]) return a[i] < m[i];
    return false;
}
}  // namespace detail

inline void init_alt_bn128_params() {}
inline void alt_bn128_seed(uint64_t seed) {
    std::lock_guard<std::mutex> l(detail::stream().mu);
    detail::stream().seed(seed);
}

struct alt_bn128_Fq {
    uint64_t v[4];
};

struct alt_bn128_Fr {
    uint64_t v[4];   // Montgomery form
    static alt_bn128_Fr zero() { return alt_bn128_Fr{{0, 0, 0, 0}}; }
    static alt_bn128_Fr one() {
        alt_bn128_Fr r;
        std::memcpy(r.v, detail::FR_ONE, 32);
        return r;
    }
    // uniform in [0, r), the value taken as the Montgomery representation
    static alt_bn128_Fr random_element() {
        auto& st = detail::stream();
        std::lock_guard<std::mutex> l(st.mu);
        alt_bn128_Fr r;
        do {
            for (auto& w : r.v) w = st.next();
            r.v[3] &= (1ULL << 62) - 1;
        } while (!detail::below(r.v, detail::FR_MOD));
        return r;
    }
    bool operator==(const alt_bn128_Fr& o) const { return std::memcmp(v, o.v, 32) == 0; }
    bool operator!=(const alt_bn128_Fr& o) const { return !(*this == o); }
    void print() const { std::printf("%016llx%016llx%016llx%016llx\n", (unsigned long long)v[3], (unsigned long long)v[2], (unsigned long long)v[1], (unsigned long long)v[0]); }
};

struct alt_bn128_G1 {
    alt_bn128_Fq X, Y, Z;   // Jacobian, Montgomery coordinates; infinity (0, 1, 0)
    static alt_bn128_G1 zero() {
        alt_bn128_G1 p{};
        std::memcpy(p.Y.v, detail::FQ_ONE, 32);
        return p;
    }
    static alt_bn128_G1 one() {
        alt_bn128_G1 p{};
        std::memcpy(p.X.v, detail::FQ_ONE, 32);
        std::memcpy(p.Y.v, detail::G1_Y, 32);
        std::memcpy(p.Z.v, detail::FQ_ONE, 32);
        return p;
    }
    void print() const {
        const alt_bn128_Fq* c[3] = {&X, &Y, &Z};
        for (auto f : c)
            std::printf("%016llx%016llx%016llx%016llx ", (unsigned long long)f->v[3], (unsigned long long)f->v[2],
                        (unsigned long long)f->v[1], (unsigned long long)f->v[0]);
        std::printf("\n");
    }
};

struct alt_bn128_G2 {
    alt_bn128_Fq X[2], Y[2], Z[2];   // Fq2 coordinates c0 | c1
    static alt_bn128_G2 zero() {
        alt_bn128_G2 p{};
        std::memcpy(p.Y[0].v, detail::FQ_ONE, 32);
        return p;
    }
    static alt_bn128_G2 one() {
        alt_bn128_G2 p{};
        std::memcpy(p.X[0].v, detail::G2_XY[0], 32);
        std::memcpy(p.X[1].v, detail::G2_XY[1], 32);
        std::memcpy(p.Y[0].v, detail::G2_XY[2], 32);
        std::memcpy(p.Y[1].v, detail::G2_XY[3], 32);
        std::memcpy(p.Z[0].v, detail::FQ_ONE, 32);
        return p;
    }
    void print() const {
        const alt_bn128_Fq* c[6] = {&X[0], &X[1], &Y[0], &Y[1], &Z[0], &Z[1]};
        for (auto f : c)
            std::printf("%016llx%016llx%016llx%016llx ", (unsigned long long)f->v[3], (unsigned long long)f->v[2],
                        (unsigned long long)f->v[1], (unsigned long long)f->v[0]);
        std::printf("\n");
    }
};

static_assert(sizeof(alt_bn128_Fr) == 32 && sizeof(alt_bn128_G1) == 96 && sizeof(alt_bn128_G2) == 192,
              "libff byte layouts");
}  // namespace libff

struct ntt_config {
    const size_t ntt_S, str_len;
    const char *forward_root_of_unity, *inverse_root_of_unity, *domain_scale_inverse;
};
inline ntt_config alt_bn128_ntt_config{28, 65, nullptr, nullptr, nullptr};
#endif  // LIGHTZK_EXTERNAL_LIBFF
