// Seeded synthetic inputs for the benchmark and serving workloads.
//
// The reference draws every input from std::random_device
// (include/host-arith/field/bigint.cuh:172-182), so no two runs agree. Here
// every input comes from a named stream:
//   key   = (seed << 32) ^ stream
//   state = four successive SplitMix64 outputs from key (Steele, Lea, Flood 2014)
//   next  = xoshiro256** (Blackman, Vigna 2018)
// A field element takes four words as little-endian limbs, clears bits >= 254
// and is rejected if >= modulus; the accepted value is used as the Montgomery
// representation (libff random_element, include/host-arith/field/fp.cuh:789-816).
// Sparse scalars: w = next(); w % 10 < 7 gives 0 or 1 (bit 32 of w), else a
// uniform element.
#include <cstdint>
#include <cstring>
#include <omp.h>
#include "../../include/hzk.h"

namespace {

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

const uint64_t FR_MOD[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
const uint64_t FQ_MOD[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
const uint64_t FR_ONE[4] = {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL};

bool below(const uint64_t* a, const uint64_t* m) {
    for (int i = 3; i >= 0; --i)
        if (a[i] != m[i]) return a[i] < m[i];
    return false;
}

void draw(Xoshiro& g, const uint64_t* mod, uint64_t* out) {
    for (;;) {
        for (int i = 0; i < 4; ++i) out[i] = g.next();
        out[3] &= (1ULL << 62) - 1;
        if (below(out, mod)) return;
    }
}

}  // namespace

extern "C" {

// kind: 0 uniform Fr, 1 sparse Fr (70% in {0,1}), 2 uniform Fq. Host output.
int hzk_gen_field(uint64_t seed, uint64_t stream, size_t n, int kind, void* host_out) {
    if (!host_out || kind < 0 || kind > 2) return HZK_ERR_INVALID;
    uint64_t* out = (uint64_t*)host_out;
    Xoshiro g(seed, stream);
    for (size_t i = 0; i < n; ++i) {
        uint64_t* o = out + 4 * i;
        if (kind == 1) {
            uint64_t w = g.next();
            if (w % 10 < 7) {
                if ((w >> 32) & 1) std::memcpy(o, FR_ONE, 32);
                else std::memset(o, 0, 32);
                continue;
            }
        }
        draw(g, kind == 2 ? FQ_MOD : FR_MOD, o);
    }
    return HZK_OK;
}

}  // extern "C"
