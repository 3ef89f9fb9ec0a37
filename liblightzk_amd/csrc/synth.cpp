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
#include "synth.hpp"

using namespace hzk::synth;

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
