// ORACLE — test infrastructure only (see orc_field.hpp). Never linked by the
// product; loaded by tests/ (ctypes), __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.
//
// C entry points over the CPU restatement: field and group operations,
// naive MSM (test/msm-check.cu:43-57 semantics), a CPU Pippenger, the
// reference NTT operator (include/operator/poly.cuh:32-73), the NTT twiddle
// table builder (include/protocol/groth16-operator.cuh:6-32), the 3-matrix
// CSR SpMV (poly.cuh:5-30) and the seeded input generators.
//
// Byte layouts at this boundary (all little-endian, Montgomery form):
//   Fr/Fq element  32 B
//   G1 affine      64 B  x|y           (infinity = all zero bytes)
//   G1 Jacobian    96 B  X|Y|Z         (normalised results: (x,y,1) or (0,1,0))
//   G2 affine     128 B  x.c0|x.c1|y.c0|y.c1 (infinity = all zero bytes)
//   G2 Jacobian   192 B  X.c0|X.c1|Y.c0|Y.c1|Z.c0|Z.c1  (alt_bn128_t.cuh:252-256)
#include <omp.h>
#include <algorithm>
#include <cstdio>
#include <vector>
#include "orc_curve.hpp"
#include "orc_rng.hpp"
#include "orc_msm.hpp"

using namespace orc;

namespace {

G1 g1_from_affine(const uint64_t* a) {
    Fq x = Fq::from_raw(a), y = Fq::from_raw(a + 4);
    if (x.is_zero() && y.is_zero()) return G1::zero();
    return {x, y, Fq::one()};
}
void g1_to_affine_bytes(const G1& P, uint64_t* out) {
    G1 A = P.to_affine();
    if (A.is_zero()) { std::memset(out, 0, 64); return; }
    std::memcpy(out, A.X.v, 32); std::memcpy(out + 4, A.Y.v, 32);
}
void g1_to_jac_bytes(const G1& P, uint64_t* out) {
    G1 A = P.to_affine();
    std::memcpy(out, A.X.v, 32); std::memcpy(out + 4, A.Y.v, 32); std::memcpy(out + 8, A.Z.v, 32);
}
G1 g1_from_jac(const uint64_t* a) {
    return {Fq::from_raw(a), Fq::from_raw(a + 4), Fq::from_raw(a + 8)};
}
Fq2 fq2_from(const uint64_t* a) { return {Fq::from_raw(a), Fq::from_raw(a + 4)}; }
void fq2_to(const Fq2& v, uint64_t* out) { std::memcpy(out, v.c0.v, 32); std::memcpy(out + 4, v.c1.v, 32); }
G2 g2_from_affine(const uint64_t* a) {
    Fq2 x = fq2_from(a), y = fq2_from(a + 8);
    if (x.is_zero() && y.is_zero()) return G2::zero();
    return {x, y, Fq2::one()};
}
void g2_to_affine_bytes(const G2& P, uint64_t* out) {
    G2 A = P.to_affine();
    if (A.is_zero()) { std::memset(out, 0, 128); return; }
    fq2_to(A.X, out); fq2_to(A.Y, out + 8);
}
void g2_to_jac_bytes(const G2& P, uint64_t* out) {
    G2 A = P.to_affine();
    fq2_to(A.X, out); fq2_to(A.Y, out + 8); fq2_to(A.Z, out + 16);
}
G2 g2_from_jac(const uint64_t* a) { return {fq2_from(a), fq2_from(a + 8), fq2_from(a + 16)}; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------- fields
// op: 0 mul, 1 add, 2 sub, 3 neg(a), 4 sqr(a), 5 inv(a), 6 to_canonical(a), 7 from_canonical(a)
int orc_fq_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        Fq x = Fq::from_raw(a + 4 * i), y = b ? Fq::from_raw(b + 4 * i) : Fq::zero(), r;
        switch (op) {
            case 0: r = x * y; break;
            case 1: r = x + y; break;
            case 2: r = x - y; break;
            case 3: r = -x; break;
            case 4: r = x.squared(); break;
            case 5: r = x.inverse(); break;
            case 6: x.to_canonical(r.v); break;
            case 7: r = Fq::from_canonical(x.v); break;
            default: return -1;
        }
        std::memcpy(out + 4 * i, r.v, 32);
    }
    return 0;
}
int orc_fr_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        Fr x = Fr::from_raw(a + 4 * i), y = b ? Fr::from_raw(b + 4 * i) : Fr::zero(), r;
        switch (op) {
            case 0: r = x * y; break;
            case 1: r = x + y; break;
            case 2: r = x - y; break;
            case 3: r = -x; break;
            case 4: r = x.squared(); break;
            case 5: r = x.inverse(); break;
            case 6: x.to_canonical(r.v); break;
            case 7: r = Fr::from_canonical(x.v); break;
            default: return -1;
        }
        std::memcpy(out + 4 * i, r.v, 32);
    }
    return 0;
}
// op: 0 mul, 1 add, 2 sub, 3 neg, 4 sqr, 5 inv
int orc_fq2_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        Fq2 x = fq2_from(a + 8 * i), y = b ? fq2_from(b + 8 * i) : Fq2::zero(), r;
        switch (op) {
            case 0: r = x * y; break;
            case 1: r = x + y; break;
            case 2: r = x - y; break;
            case 3: r = -x; break;
            case 4: r = x.squared(); break;
            case 5: r = x.inverse(); break;
            default: return -1;
        }
        fq2_to(r, out + 8 * i);
    }
    return 0;
}

// ---------------------------------------------------------------- groups
// op: 0 add(a,b), 1 dbl(a), 2 neg(a), 3 normalise(a); inputs/outputs Jacobian
int orc_g1_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        G1 x = g1_from_jac(a + 12 * i), r;
        switch (op) {
            case 0: r = x + g1_from_jac(b + 12 * i); break;
            case 1: r = x.dbl(); break;
            case 2: r = x.neg(); break;
            case 3: r = x; break;
            default: return -1;
        }
        g1_to_jac_bytes(r, out + 12 * i);
    }
    return 0;
}
int orc_g2_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        G2 x = g2_from_jac(a + 24 * i), r;
        switch (op) {
            case 0: r = x + g2_from_jac(b + 24 * i); break;
            case 1: r = x.dbl(); break;
            case 2: r = x.neg(); break;
            case 3: r = x; break;
            default: return -1;
        }
        g2_to_jac_bytes(r, out + 24 * i);
    }
    return 0;
}
int orc_g1_on_curve(const uint64_t* jac) { return on_curve_g1(g1_from_jac(jac)) ? 1 : 0; }
int orc_g2_on_curve(const uint64_t* jac) { return on_curve_g2(g2_from_jac(jac)) ? 1 : 0; }
void orc_g1_generator(uint64_t* out_jac) { g1_to_jac_bytes(g1_generator(), out_jac); }
void orc_g2_generator(uint64_t* out_jac) { g2_to_jac_bytes(g2_generator(), out_jac); }

// k (Fr, Montgomery) times the generator, affine output (fixed-base comb +
// batch normalisation; the scalar_mul path is orc_g*_scalar_mul below).
int orc_g1_mul_gen(const uint64_t* k, uint64_t* out_aff, size_t n) {
    static FixedBase<G1> fb(g1_generator());
    std::vector<G1> pts(n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) {
        uint64_t c[4]; Fr::from_raw(k + 4 * i).to_canonical(c);
        pts[i] = fb.mul(c);
    }
    batch_to_affine<G1, Fq>(pts);
    for (size_t i = 0; i < n; ++i) g1_to_affine_bytes(pts[i], out_aff + 8 * i);
    return 0;
}
int orc_g2_mul_gen(const uint64_t* k, uint64_t* out_aff, size_t n) {
    static FixedBase<G2> fb(g2_generator());
    std::vector<G2> pts(n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) {
        uint64_t c[4]; Fr::from_raw(k + 4 * i).to_canonical(c);
        pts[i] = fb.mul(c);
    }
    batch_to_affine<G2, Fq2>(pts);
    for (size_t i = 0; i < n; ++i) g2_to_affine_bytes(pts[i], out_aff + 16 * i);
    return 0;
}
// scalar (Fr, Montgomery) times an affine point by MSB-first double-and-add
// (curve_utils.cuh:29-50); Jacobian normalised output.
int orc_g1_scalar_mul(const uint64_t* k, const uint64_t* p_aff, uint64_t* out_jac) {
    uint64_t c[4]; Fr::from_raw(k).to_canonical(c);
    g1_to_jac_bytes(g1_from_affine(p_aff).scalar_mul(c), out_jac);
    return 0;
}
int orc_g2_scalar_mul(const uint64_t* k, const uint64_t* p_aff, uint64_t* out_jac) {
    uint64_t c[4]; Fr::from_raw(k).to_canonical(c);
    g2_to_jac_bytes(g2_from_affine(p_aff).scalar_mul(c), out_jac);
    return 0;
}

// ---------------------------------------------------------------- MSM
// method: 0 naive (per-thread s_i*P_i accumulators, test/msm-check.cu:46-55), 1 CPU Pippenger
// Output: Jacobian normalised (96 B / 192 B).
int orc_msm_g1(const uint64_t* scalars, const uint64_t* bases_aff, size_t n, int method, uint64_t* out_jac) {
    auto fa = [bases_aff](size_t i) { return g1_from_affine(bases_aff + 8 * i); };
    G1 r = method == 0 ? msm_naive<G1>(scalars, n, fa) : msm_pippenger<G1>(scalars, n, fa);
    g1_to_jac_bytes(r, out_jac);
    return 0;
}
int orc_msm_g2(const uint64_t* scalars, const uint64_t* bases_aff, size_t n, int method, uint64_t* out_jac) {
    auto fa = [bases_aff](size_t i) { return g2_from_affine(bases_aff + 16 * i); };
    G2 r = method == 0 ? msm_naive<G2>(scalars, n, fa) : msm_pippenger<G2>(scalars, n, fa);
    g2_to_jac_bytes(r, out_jac);
    return 0;
}
// Closed form for bases P_i = k_i*G: sum s_i*P_i = (sum s_i*k_i mod r)*G.
int orc_msm_closed_form(const uint64_t* scalars, const uint64_t* ks, size_t n, int group, uint64_t* out_jac) {
    int nt = omp_get_max_threads();
    std::vector<Fr> part(nt, Fr::zero());
#pragma omp parallel
    {
        int t = omp_get_thread_num();
#pragma omp for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            // s (Montgomery) * canonical(k): the product of two Montgomery values
            // is Montgomery(s*k); sum in Montgomery form, convert at the end.
            part[t] += Fr::from_raw(scalars + 4 * i) * Fr::from_raw(ks + 4 * i);
        }
    }
    Fr acc = Fr::zero();
    for (auto& v : part) acc += v;
    uint64_t c[4]; acc.to_canonical(c);
    if (group == 1) g1_to_jac_bytes(g1_generator().scalar_mul(c), out_jac);
    else g2_to_jac_bytes(g2_generator().scalar_mul(c), out_jac);
    return 0;
}

// ---------------------------------------------------------------- generators
// kind: 0 uniform; 1 sparse (70% in {0,1}, 30% uniform)
int orc_gen_fr(uint64_t seed, uint64_t stream, size_t n, int kind, uint64_t* out) {
    Rng rng(seed, stream);
    for (size_t i = 0; i < n; ++i) {
        Fr v;
        if (kind == 1) {
            uint64_t w = rng.next();
            if (w % 10 < 7) v = ((w >> 32) & 1) ? Fr::one() : Fr::zero();
            else v = rng.field<Fr>();
        } else {
            v = rng.field<Fr>();
        }
        std::memcpy(out + 4 * i, v.v, 32);
    }
    return 0;
}
int orc_gen_fq(uint64_t seed, uint64_t stream, size_t n, uint64_t* out) {
    Rng rng(seed, stream);
    for (size_t i = 0; i < n; ++i) { Fq v = rng.field<Fq>(); std::memcpy(out + 4 * i, v.v, 32); }
    return 0;
}
uint64_t orc_rng_u64(uint64_t seed, uint64_t stream, size_t n, uint64_t* out) {
    Rng rng(seed, stream);
    for (size_t i = 0; i < n; ++i) out[i] = rng.next();
    return n;
}

// ---------------------------------------------------------------- NTT
// The reference operator (poly.cuh:32-73): DIF radix-2 stages with twiddle
// omega_powers[(j & (hl-1)) << i], bit-reverse permutation, optional scale.
// Result: out[k] = factor * sum_j in[j] * w^(jk), w = omega_powers[1].
int orc_ntt(size_t deg, size_t log_deg, uint64_t* coeffs, const uint64_t* omega_powers, const uint64_t* factor) {
    if (deg != ((size_t)1 << log_deg)) return -1;
    Fr* a = reinterpret_cast<Fr*>(coeffs);
    const Fr* w = reinterpret_cast<const Fr*>(omega_powers);
    for (size_t hl = deg >> 1, i = 0; hl > 0; hl >>= 1, ++i) {
#pragma omp parallel for schedule(static) if (deg >= 4096)
        for (long jj = 0; jj < (long)deg; ++jj) {
            size_t j = (size_t)jj;
            if (j & hl) continue;
            Fr lo = a[j], hi = a[j ^ hl];
            a[j] = lo + hi;
            a[j ^ hl] = (lo - hi) * w[(j & (hl - 1)) << i];
        }
    }
    for (size_t j = 0; j < deg; ++j) {
        size_t r = 0;
        for (size_t b = 0; b < log_deg; ++b) r |= ((j >> b) & 1) << (log_deg - 1 - b);
        if (j < r) std::swap(a[j], a[r]);
    }
    if (factor) {
        Fr f = Fr::from_raw(factor);
        for (size_t j = 0; j < deg; ++j) a[j] *= f;
    }
    return 0;
}

// Primitive 2^k-th root of unity in Fr: 5^((r-1)/2^k) (multiplicative
// generator 5, alt_bn128_init.cu:50); k <= 28.
int orc_fr_root_of_unity(int k, int inverse, uint64_t* out) {
    if (k < 0 || k > 28) return -1;
    uint64_t e[4]; const uint64_t one_raw[4] = {1, 0, 0, 0};
    sub4(e, FR_PARAMS.p, one_raw);      // r - 1
    for (int s = 0; s < k; ++s) {       // >> k
        for (int i = 0; i < 4; ++i) e[i] = (e[i] >> 1) | (i < 3 ? e[i + 1] << 63 : 0);
    }
    Fr w = Fr::from_u64(5).pow(e, 4);
    if (inverse) w = w.inverse();
    std::memcpy(out, w.v, 32);
    return 0;
}

// Twiddle table of include/protocol/groth16-operator.cuh:6-32 (6M+2 entries):
// [w_M^i]_M [w_M^-i]_M [M^-1] [w_2M^i]_2M [w_2M^-i]_2M [(4M)^-1]
int orc_ntt_setup(size_t M, uint64_t* out) {
    int lg = 0; while (((size_t)1 << lg) < M) ++lg;
    if (((size_t)1 << lg) != M || lg + 2 > 28) return -1;
    Fr* t = reinterpret_cast<Fr*>(out);
    Fr wM, wMi, w2M, w2Mi;
    orc_fr_root_of_unity(lg, 0, wM.v); orc_fr_root_of_unity(lg, 1, wMi.v);
    orc_fr_root_of_unity(lg + 1, 0, w2M.v); orc_fr_root_of_unity(lg + 1, 1, w2Mi.v);
    Fr two_inv = Fr::from_u64(2).inverse();
    auto pow2inv = [&](int k) { Fr r = Fr::one(); for (int i = 0; i < k; ++i) r *= two_inv; return r; };
    t[0] = t[M] = t[2 * M + 1] = t[4 * M + 1] = Fr::one();
    t[2 * M] = pow2inv(lg);
    t[6 * M + 1] = pow2inv(lg + 2);
    for (size_t i = 1; i < M; ++i) { t[i] = t[i - 1] * wM; t[M + i] = t[M + i - 1] * wMi; }
    for (size_t i = 1; i < 2 * M; ++i) { t[2 * M + 1 + i] = t[2 * M + i] * w2M; t[4 * M + 1 + i] = t[4 * M + i] * w2Mi; }
    return 0;
}

// ---------------------------------------------------------------- SpMV
// The reference's CSR (include/data-structure/xpu-spmat.cuh:20-25): row_ptr has
// `row` entries and no sentinel; the last row ends at cnt (poly.cuh:18-29).
int orc_spmv(size_t row, size_t cnt, const uint64_t* val, const uint64_t* col_idx, const uint64_t* row_ptr,
             const uint64_t* z, uint64_t* out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)row; ++i) {
        size_t b = row_ptr[i], e = (size_t)i + 1 < row ? row_ptr[i + 1] : cnt;
        Fr acc = Fr::zero();
        for (size_t j = b; j < e; ++j) acc += Fr::from_raw(z + 4 * col_idx[j]) * Fr::from_raw(val + 4 * j);
        std::memcpy(out + 4 * i, acc.v, 32);
    }
    return 0;
}

int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
