// ORACLE — test infrastructure only (see orc_field.hpp).
//
// CPU multi-scalar multiplication and fixed-base scalar multiplication used by
// the oracle entry points and the oracle's Groth16 prover.
#pragma once
#include <omp.h>
#include <vector>
#include "orc_curve.hpp"

namespace orc {

// get_base(i) returns base i as a Jacobian point (Z = 1 or zero).
template <class G, class GetBase>
G msm_naive(const uint64_t* scalars, size_t n, GetBase get_base) {
    int nt = omp_get_max_threads();
    std::vector<G> acc(nt, G::zero());
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
#pragma omp for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            uint64_t k[4];
            Fr::from_raw(scalars + 4 * i).to_canonical(k);
            acc[tid] = acc[tid] + get_base((size_t)i).scalar_mul(k);
        }
    }
    G r = G::zero();
    for (int t = 0; t < nt; ++t) r = r + acc[t];
    return r;
}

// Unsigned-window Pippenger on the CPU (not the reference's method; a faster
// checker): windows of c bits over the canonical scalar, buckets summed with
// mixed additions, running-sum bucket reduction, Horner window combine.
template <class G, class GetBase>
G msm_pippenger(const uint64_t* scalars, size_t n, GetBase get_base) {
    int c = 4;
    while (c < 16 && ((size_t)1 << (c + 2)) < n) ++c;
    int W = (254 + c - 1) / c;
    std::vector<uint64_t> canon(4 * n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) Fr::from_raw(scalars + 4 * i).to_canonical(&canon[4 * i]);
    std::vector<G> window_sum(W, G::zero());
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < W; ++w) {
        std::vector<G> buckets((size_t)1 << c, G::zero());
        int lo = w * c;
        for (size_t i = 0; i < n; ++i) {
            const uint64_t* k = &canon[4 * i];
            uint64_t d = 0;
            for (int b = 0; b < c && lo + b < 256; ++b) d |= ((k[(lo + b) / 64] >> ((lo + b) % 64)) & 1) << b;
            if (d) buckets[d] = buckets[d].mixed_add(get_base(i));
        }
        G run = G::zero(), sum = G::zero();
        for (size_t b = buckets.size() - 1; b >= 1; --b) { run = run + buckets[b]; sum = sum + run; }
        window_sum[w] = sum;
    }
    G r = G::zero();
    for (int w = W - 1; w >= 0; --w) {
        for (int i = 0; i < c; ++i) r = r.dbl();
        r = r + window_sum[w];
    }
    return r;
}


// Fixed-base comb for k*G: T[j][d] = d * 2^(8j) * G (32 windows x 256 digits);
// k*G = sum_j T[j][byte_j(k)] with mixed additions, then batch normalisation
// (Montgomery's trick) of the outputs.
template <class G>
struct FixedBase {
    std::vector<G> table;   // 32 * 256 affine points (Z = 1, or zero)
    explicit FixedBase(const G& base) : table(32 * 256) {
        G b = base;
        for (int j = 0; j < 32; ++j) {
            G acc = G::zero();
            table[j * 256] = G::zero();
            for (int d = 1; d < 256; ++d) { acc = acc + b; table[j * 256 + d] = acc.to_affine(); }
            for (int i = 0; i < 8; ++i) b = b.dbl();
        }
    }
    // canonical scalars in, Jacobian out
    G mul(const uint64_t k[4]) const {
        G acc = G::zero();
        for (int j = 0; j < 32; ++j) {
            unsigned d = (k[j / 8] >> (8 * (j % 8))) & 0xff;
            if (d) acc = acc.mixed_add(table[j * 256 + d]);
        }
        return acc;
    }
};

// Batch normalisation: out[i] = in[i].to_affine() with one inversion per chunk.
template <class G, class F>
void batch_to_affine(std::vector<G>& pts) {
    const size_t n = pts.size();
    const size_t chunk = 4096;
#pragma omp parallel for schedule(dynamic, 1)
    for (long c0 = 0; c0 < (long)n; c0 += chunk) {
        size_t c1 = std::min(n, (size_t)c0 + chunk);
        std::vector<F> pref(c1 - c0);
        F acc = F::one();
        for (size_t i = c0; i < c1; ++i) {
            pref[i - c0] = acc;
            if (!pts[i].is_zero()) acc = acc * pts[i].Z;
        }
        F inv = acc.inverse();
        for (size_t i = c1; i-- > (size_t)c0;) {
            if (pts[i].is_zero()) { pts[i] = G::zero(); continue; }
            F zi = inv * pref[i - c0];
            inv = inv * pts[i].Z;
            F zi2 = zi.squared();
            pts[i] = {pts[i].X * zi2, pts[i].Y * (zi2 * zi), F::one()};
        }
    }
}

}  // namespace orc
