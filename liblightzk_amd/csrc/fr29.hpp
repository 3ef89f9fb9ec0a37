// Fr in nine 29-bit limbs (fr29), for the NTT's latency-bound two-element
// passes (ntt.hip). Needs HZK_DEV and fr_t (eight 32-bit limbs, field.hpp)
// declared first; plain C++ otherwise, so the host sanitizer build checks it
// against the oracle's Fr product.
#pragma once
#include <cstdint>

namespace hzk {

// Fr in nine 29-bit limbs for the latency-bound two-element passes. The
// boundary form's Montgomery product (field.hpp: eight 32-bit limbs, one
// chain of 128 multiply-adds with carry folds) is a long dependency chain;
// here the 17 product columns are independent sums (a 64-bit column cannot
// overflow), and only the nine reduction digits form a chain of a few
// instructions each, so one product finishes in a fraction of the time.
// Montgomery radix 2^261: a product of values in the library's 2^256 form
// stays in that form when one factor (a twiddle or table entry, canonical)
// is first multiplied by 2^5. Values stay below 2r between operations.
struct fr29 {
    static constexpr uint32_t MASK = (1u << 29) - 1;
    static constexpr uint32_t P[9] = {0x10000001, 0x1f0fac9f, 0x0e5c2450, 0x07d090f3, 0x1585d283,
                                      0x02db40c0, 0x00a6e141, 0x0e5c2634, 0x0030644e};   // r
    static constexpr uint32_t P2[9] = {0x00000002, 0x1e1f593f, 0x1cb848a1, 0x0fa121e6, 0x0b0ba506,
                                       0x05b68181, 0x014dc282, 0x1cb84c68, 0x0060c89c};  // 2r
    static constexpr uint32_t M0 = 0x0fffffff;   // -r^-1 mod 2^29
    uint32_t l[9];

    // 8 x 32-bit limbs (a value < 2^256) -> 9 x 29-bit limbs
    static HZK_DEV fr29 unpack(const fr_t& x) {
        const uint32_t* w = x.l;
        fr29 r;
        r.l[0] = w[0] & MASK;
        r.l[1] = ((w[0] >> 29) | (w[1] << 3)) & MASK;
        r.l[2] = ((w[1] >> 26) | (w[2] << 6)) & MASK;
        r.l[3] = ((w[2] >> 23) | (w[3] << 9)) & MASK;
        r.l[4] = ((w[3] >> 20) | (w[4] << 12)) & MASK;
        r.l[5] = ((w[4] >> 17) | (w[5] << 15)) & MASK;
        r.l[6] = ((w[5] >> 14) | (w[6] << 18)) & MASK;
        r.l[7] = ((w[6] >> 11) | (w[7] << 21)) & MASK;
        r.l[8] = w[7] >> 8;
        return r;
    }
    // normalized limbs of a value < 2^256 -> 8 x 32-bit limbs
    HZK_DEV fr_t pack() const {
        fr_t x;
        x.l[0] = l[0] | (l[1] << 29);
        x.l[1] = (l[1] >> 3) | (l[2] << 26);
        x.l[2] = (l[2] >> 6) | (l[3] << 23);
        x.l[3] = (l[3] >> 9) | (l[4] << 20);
        x.l[4] = (l[4] >> 12) | (l[5] << 17);
        x.l[5] = (l[5] >> 15) | (l[6] << 14);
        x.l[6] = (l[6] >> 18) | (l[7] << 11);
        x.l[7] = (l[7] >> 21) | (l[8] << 8);
        return x;
    }
    // 32 x (a value < 2^256): a 2^256-form factor made a 2^261-form one
    HZK_DEV fr29 times32() const {
        fr29 r;
        r.l[0] = (l[0] << 5) & MASK;
#pragma unroll
        for (int i = 1; i < 9; ++i) r.l[i] = ((l[i] << 5) | (l[i - 1] >> 24)) & MASK;
        return r;
    }
    // a + b, normalized (a + b < 2^256)
    static HZK_DEV fr29 add(const fr29& a, const fr29& b) {
        fr29 r;
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const uint32_t v = a.l[i] + b.l[i] + c;
            r.l[i] = v & MASK;
            c = v >> 29;
        }
        return r;
    }
    // a - b + 2r for b <= 2r, normalized
    static HZK_DEV fr29 sub2(const fr29& a, const fr29& b) {
        fr29 r;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int32_t v = (int32_t)(a.l[i] + P2[i] - b.l[i]) + c;
            r.l[i] = (uint32_t)v & MASK;
            c = v >> 29;
        }
        return r;
    }
    // a normalized value below 2^261 brought below 2r: subtract q r with
    // q = floor(l8 / (r8 + 1)) <= floor(value / r); what remains is below
    // (r8 + 1 + q) 2^232 < 2r
    HZK_DEV fr29 reduce() const {
        const uint32_t q = l[8] / (P[8] + 1);
        fr29 r;
        int64_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int64_t v = (int64_t)l[i] - (int64_t)q * P[i] + c;
            r.l[i] = (uint32_t)v & MASK;
            c = v >> 29;
        }
        return r;
    }
    // a value below 2r -> below r
    HZK_DEV fr29 canonical() const {
        fr29 d;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int32_t v = (int32_t)l[i] - (int32_t)P[i] + c;
            d.l[i] = (uint32_t)v & MASK;
            c = v >> 29;
        }
        return c < 0 ? *this : d;
    }
    // a b 2^-261 mod r, below 2r when the bounds of a and b (in units of r)
    // multiply to less than 2^261 / r > 169; normalized inputs. A column holds
    // at most 9 limb products and 9 reduction products (< 2^58 each) and a
    // carry: < 2^62.2.
    static HZK_DEV fr29 mul(const fr29& a, const fr29& b) {
        uint64_t c[18];
#pragma unroll
        for (int k = 0; k < 17; ++k) {
            uint64_t s = 0;
#pragma unroll
            for (int j = (k > 8 ? k - 8 : 0); j <= (k < 8 ? k : 8); ++j) s += (uint64_t)a.l[j] * b.l[k - j];
            c[k] = s;
        }
        c[17] = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const uint32_t m = ((uint32_t)c[k] * M0) & MASK;
#pragma unroll
            for (int i = 0; i < 9; ++i) c[k + i] += (uint64_t)m * P[i];
            c[k + 1] += c[k] >> 29;
        }
        fr29 r;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            r.l[i] = (uint32_t)c[9 + i] & MASK;
            c[10 + i] += c[9 + i] >> 29;
        }
        r.l[8] = (uint32_t)c[17];
        return r;
    }
};

}  // namespace hzk
