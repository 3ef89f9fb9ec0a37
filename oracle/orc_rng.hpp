// ORACLE — test infrastructure only (see orc_field.hpp).
//
// Seeded synthetic-input generator. The reference draws every input from
// std::random_device (include/host-arith/field/bigint.cuh:172-182), so no run
// repeats; this build replaces that with a seeded stream that both the CPU
// oracle and the GPU path consume byte-for-byte:
//   key   = (seed << 32) ^ stream
//   state = four successive SplitMix64 outputs starting from key
//           (Steele, Lea, Flood 2014)
//   next  = xoshiro256** (Blackman, Vigna 2018)
// A field element draws four words (little-endian limbs), clears every bit at
// or above bit 254 and rejects values >= modulus; the accepted value IS the
// Montgomery representation, as libff's Fp_model::random_element does
// (include/host-arith/field/fp.cuh:789-816).
#pragma once
#include <cstdint>
#include "orc_field.hpp"

namespace orc {

struct Rng {
    uint64_t s[4];

    static uint64_t splitmix64(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = (seed << 32) ^ stream;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t result = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    template <class F>
    F field() {
        const uint64_t* p = F::modulus();
        for (;;) {
            uint64_t w[4];
            for (int i = 0; i < 4; ++i) w[i] = next();
            w[3] &= (1ULL << 62) - 1;   // keep bits 0..253
            if (cmp4(w, p) < 0) return F::from_raw(w);
        }
    }
};

}  // namespace orc
