// Seeded synthetic-input streams (see synth.cpp for the specification).
#pragma once
#include <cstdint>
#include <cstring>

namespace hzk {
namespace synth {


struct Xoshiro {
    uint64_t s[4];
    Xoshiro(uint64_t seed, uint64_t stream) {
        uint64_t x = (seed << 32) ^ stream;
        for (int i = 0; i < 4; ++i) {
            uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
};

inline constexpr uint64_t FR_MOD[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
inline constexpr uint64_t FQ_MOD[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
inline constexpr uint64_t FR_ONE[4] = {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL};

inline bool below(const uint64_t* a, const uint64_t* m) {
    for (int i = 3; i >= 0; --i)
        if (a[i] != m[i]) return a[i] < m[i];
    return false;
}

inline void draw(Xoshiro& g, const uint64_t* mod, uint64_t* out) {
    for (;;) {
        for (int i = 0; i < 4; ++i) out[i] = g.next();
        out[3] &= (1ULL << 62) - 1;
        if (below(out, mod)) return;
    }
}

}  // namespace synth
}  // namespace hzk
