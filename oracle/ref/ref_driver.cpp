// ORACLE — test infrastructure only. C entry points over the reference's own
// host arithmetic (the libff subset vendored under include/host-arith and
// src/host-arith of the reference tree), compiled from those sources by
// oracle/ref/Makefile into oracle/_ref/libref_hostarith.so. Used to check the
// oracle's restatement and as the "reference" CPU baseline; never linked by
// the product. Buffers use the same 32-byte Montgomery layout as the oracle
// (libff Fp_model::mont_repr, include/host-arith/field/fp.cuh:42).
#include <omp.h>
#include <cstring>
#include <vector>
#include "alt_bn128_g1.cuh"
#include "alt_bn128_g2.cuh"
#include "alt_bn128_init.cuh"

using namespace libff;

namespace {

alt_bn128_Fq fq(const uint64_t* p) { alt_bn128_Fq r; std::memcpy(r.mont_repr.data, p, 32); return r; }
alt_bn128_Fr fr(const uint64_t* p) { alt_bn128_Fr r; std::memcpy(r.mont_repr.data, p, 32); return r; }
void put(const alt_bn128_Fq& v, uint64_t* o) { std::memcpy(o, v.mont_repr.data, 32); }
alt_bn128_Fq2 fq2(const uint64_t* p) { return alt_bn128_Fq2(fq(p), fq(p + 4)); }
void put2(const alt_bn128_Fq2& v, uint64_t* o) { put(v.c0, o); put(v.c1, o + 4); }

alt_bn128_G1 g1_aff(const uint64_t* a) {
    alt_bn128_Fq x = fq(a), y = fq(a + 4);
    if (x.is_zero() && y.is_zero()) return alt_bn128_G1::zero();
    return alt_bn128_G1(x, y, alt_bn128_Fq::one());
}
alt_bn128_G2 g2_aff(const uint64_t* a) {
    alt_bn128_Fq2 x = fq2(a), y = fq2(a + 8);
    if (x.is_zero() && y.is_zero()) return alt_bn128_G2::zero();
    return alt_bn128_G2(x, y, alt_bn128_Fq2::one());
}
// normalised Jacobian out (libff to_affine_coordinates, inf -> (0, 1, 0))
void put_g1(alt_bn128_G1 p, uint64_t* o) { p.to_affine_coordinates(); put(p.X, o); put(p.Y, o + 4); put(p.Z, o + 8); }
void put_g2(alt_bn128_G2 p, uint64_t* o) { p.to_affine_coordinates(); put2(p.X, o); put2(p.Y, o + 8); put2(p.Z, o + 16); }

}  // namespace

extern "C" {

void ref_init() {
    static bool done = false;
    if (!done) { init_alt_bn128_params(); done = true; }
}

// op: 0 mul, 1 add, 2 sub, 3 neg, 4 sqr, 5 inv
int ref_fq_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) {
        alt_bn128_Fq x = fq(a + 4 * i), y = b ? fq(b + 4 * i) : alt_bn128_Fq::zero(), r;
        switch (op) {
            case 0: r = x * y; break; case 1: r = x + y; break; case 2: r = x - y; break;
            case 3: r = -x; break; case 4: r = x.squared(); break; case 5: r = x.is_zero() ? x : x.inverse(); break;
            default: return -1;
        }
        put(r, out + 4 * i);
    }
    return 0;
}
int ref_fr_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) {
        alt_bn128_Fr x = fr(a + 4 * i), y = b ? fr(b + 4 * i) : alt_bn128_Fr::zero(), r;
        switch (op) {
            case 0: r = x * y; break; case 1: r = x + y; break; case 2: r = x - y; break;
            case 3: r = -x; break; case 4: r = x.squared(); break; case 5: r = x.is_zero() ? x : x.inverse(); break;
            default: return -1;
        }
        std::memcpy(out + 4 * i, r.mont_repr.data, 32);
    }
    return 0;
}
int ref_fq2_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) {
        alt_bn128_Fq2 x = fq2(a + 8 * i), y = b ? fq2(b + 8 * i) : alt_bn128_Fq2::zero(), r;
        switch (op) {
            case 0: r = x * y; break; case 1: r = x + y; break; case 2: r = x - y; break;
            case 3: r = -x; break; case 4: r = x.squared(); break; case 5: r = x.is_zero() ? x : x.inverse(); break;
            default: return -1;
        }
        put2(r, out + 8 * i);
    }
    return 0;
}
// k (Fr Montgomery) * affine point, normalised Jacobian out
int ref_g1_scalar_mul(const uint64_t* k, const uint64_t* p_aff, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) put_g1(fr(k + 4 * i) * g1_aff(p_aff + 8 * i), out + 12 * i);
    return 0;
}
int ref_g2_scalar_mul(const uint64_t* k, const uint64_t* p_aff, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) put_g2(fr(k + 4 * i) * g2_aff(p_aff + 16 * i), out + 24 * i);
    return 0;
}
// a + b over affine inputs (exercises add-2007-bl incl. doubling / inverse cases)
int ref_g1_add(const uint64_t* a_aff, const uint64_t* b_aff, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) put_g1(g1_aff(a_aff + 8 * i) + g1_aff(b_aff + 8 * i), out + 12 * i);
    return 0;
}
int ref_g2_add(const uint64_t* a_aff, const uint64_t* b_aff, uint64_t* out, size_t n) {
    ref_init();
    for (size_t i = 0; i < n; ++i) put_g2(g2_aff(a_aff + 16 * i) + g2_aff(b_aff + 16 * i), out + 24 * i);
    return 0;
}
// test/msm-check.cu:46-55: per-thread accumulators of s_i * P_i, then a serial reduce
int ref_msm_g1(const uint64_t* s, const uint64_t* bases_aff, size_t n, uint64_t* out) {
    ref_init();
    int nt = omp_get_max_threads();
    std::vector<alt_bn128_G1> acc(nt, alt_bn128_G1::zero());
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
#pragma omp for
        for (long i = 0; i < (long)n; ++i) acc[tid] = acc[tid] + fr(s + 4 * i) * g1_aff(bases_aff + 8 * i);
    }
    for (int t = 1; t < nt; ++t) acc[0] = acc[0] + acc[t];
    put_g1(acc[0], out);
    return 0;
}
int ref_msm_g2(const uint64_t* s, const uint64_t* bases_aff, size_t n, uint64_t* out) {
    ref_init();
    int nt = omp_get_max_threads();
    std::vector<alt_bn128_G2> acc(nt, alt_bn128_G2::zero());
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
#pragma omp for
        for (long i = 0; i < (long)n; ++i) acc[tid] = acc[tid] + fr(s + 4 * i) * g2_aff(bases_aff + 16 * i);
    }
    for (int t = 1; t < nt; ++t) acc[0] = acc[0] + acc[t];
    put_g2(acc[0], out);
    return 0;
}
int ref_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
