// ORACLE — test infrastructure only (see orc_field.hpp).
//
// Groth16 over alt_bn128 on the CPU: the synthetic multiplication-gate
// circuit, QAP key generation from seeded toxic waste, the honest prover
// (phase order of include/protocol/groth16.cuh:154-222 with libsnark-style
// H: 3 iNTT + 3 coset NTT + 1 coset iNTT), a trapdoor closed-form prover that
// needs no MSM, and the pairing verifier.
//
// The reference's Groth16 is a timing skeleton (random keys, random witness,
// no verifier: groth16.cuh:84-132), so this restates the protocol itself:
//   QAP: rows j < M on the domain {w^j}, L_j(x) = (x^M - 1) w^j / (M (x - w^j)),
//        A_i(tau) = sum_j A[j][i] L_j(tau) (same for B, C); the last
//        num_public+1 rows hold z_i in A only (input consistency rows).
//   pk:  A_i(tau) G1, B_i(tau) G1, B_i(tau) G2 (all i),
//        (beta A_i + alpha B_i + C_i)/delta G1 (i > num_public),
//        tau^j Z(tau)/delta G1 (j < M-1),
//        alpha G1, beta G1, delta G1, beta G2, delta G2
//   vk:  alpha G1, beta G2, gamma G2, delta G2, IC_i = (beta A_i + alpha B_i + C_i)/gamma G1 (i <= num_public)
//   proof (phase 6 of groth16.cuh:210-215):
//        A = alpha1 + sum z_i A_i + r delta1
//        B = beta2 + sum z_i B2_i + s delta2,  B1 = beta1 + sum z_i B1_i + s delta1
//        C = sum_{i>l} z_i L_i + sum_j h_j H_j + s A + r B1 - r s delta1
//   verify: e(A,B) = e(alpha1,beta2) e(sum_{i<=l} z_i IC_i, gamma2) e(C, delta2)
#include <cstdio>
#include <vector>
#include "orc_curve.hpp"
#include "orc_msm.hpp"
#include "orc_rng.hpp"

using namespace orc;

extern "C" int orc_ntt(size_t deg, size_t log_deg, uint64_t* coeffs, const uint64_t* omega_powers, const uint64_t* factor);
extern "C" int orc_fr_root_of_unity(int k, int inverse, uint64_t* out);
extern "C" int orc_pairing_product_is_one(const uint64_t* ps_aff, const uint64_t* qs_aff, size_t n);

namespace {

const int NUM_INPUTS = 64;   // seeded witness inputs z[1..64]; the first num_public are public

struct Csr {
    std::vector<uint64_t> row_ptr, col;   // row_ptr has `rows` entries, no sentinel
    std::vector<Fr> val;
};

struct Circuit {
    int log_M;
    int kind = 0;   // 0 multiplication-gate DAG, 1 MatMul (p = mm_p)
    int mm_p = 0;
    size_t M, num_vars, num_public, num_gates;
    Csr A, B, C;
};

Circuit make_circuit(uint64_t seed, int log_M, int num_public) {
    Circuit c;
    c.log_M = log_M;
    c.M = (size_t)1 << log_M;
    c.num_public = num_public;
    c.num_gates = c.M - (num_public + 1);
    c.num_vars = 1 + NUM_INPUTS + c.num_gates;
    Rng rng(seed, 2);
    for (size_t g = 0; g < c.num_gates; ++g) {
        uint64_t v = 1 + NUM_INPUTS + g;
        Csr* mats[2] = {&c.A, &c.B};
        for (Csr* m : mats) {
            m->row_ptr.push_back(m->col.size());
            int k = (int)(rng.next() % 3) + 1;
            for (int t = 0; t < k; ++t) {
                m->col.push_back(rng.next() % v);
                m->val.push_back(rng.field<Fr>());
            }
        }
        c.C.row_ptr.push_back(c.C.col.size());
        c.C.col.push_back(v);
        c.C.val.push_back(Fr::one());
    }
    for (size_t i = 0; i <= (size_t)num_public; ++i) {
        c.A.row_ptr.push_back(c.A.col.size());
        c.A.col.push_back(i);
        c.A.val.push_back(Fr::one());
        c.B.row_ptr.push_back(c.B.col.size());
        c.C.row_ptr.push_back(c.C.col.size());
    }
    return c;
}

// The reference's MatMul R1CS (include/data-structure/xpu-spmat.cuh:33-65,
// sizes include/protocol/groth16.cuh:11-16): p^3 + p^2 constraint rows padded
// with empty rows to M = next_pow2(p^3 + p^2), N + 1 = 1 + 3p^2 + p^3
// variables, every coefficient one. Row i < p^3 with (i1, i2, i3) the base-p
// digits of i: A = z[1 + i2 p + i1], B = z[1 + p^2 + i1 p + i3],
// C = z[1 + 3p^2 + i]. Row p^3 + i, i < p^2: A = sum_k z[1 + 3p^2 + k p^2 + i],
// B = z[0], C = z[1 + 2p^2 + i]. Public inputs: the first k = 3p^2 variables.
Circuit make_circuit_matmul(int p) {
    Circuit c;
    const size_t P = (size_t)p, p2 = P * P, p3 = p2 * P;
    c.kind = 1;
    c.mm_p = p;
    c.log_M = 0;
    while (((size_t)1 << c.log_M) < p3 + p2) ++c.log_M;
    if (c.log_M < 4) c.log_M = 4;
    c.M = (size_t)1 << c.log_M;
    c.num_public = 3 * p2;
    c.num_gates = p3 + p2;
    c.num_vars = 1 + 3 * p2 + p3;
    auto row = [&](Csr& m) { m.row_ptr.push_back(m.col.size()); };
    auto one = [&](Csr& m, uint64_t col) { m.col.push_back(col); m.val.push_back(Fr::one()); };
    for (size_t i = 0; i < p3; ++i) {
        size_t i1 = i / p2, i2 = (i / P) % P, i3 = i % P;
        row(c.A); one(c.A, 1 + i2 * P + i1);
        row(c.B); one(c.B, 1 + p2 + i1 * P + i3);
        row(c.C); one(c.C, 1 + 3 * p2 + i);
    }
    for (size_t i = 0; i < p2; ++i) {
        row(c.A);
        for (size_t k = 0; k < P; ++k) one(c.A, 1 + 3 * p2 + k * p2 + i);
        row(c.B); one(c.B, 0);
        row(c.C); one(c.C, 1 + 2 * p2 + i);
    }
    for (size_t r = p3 + p2; r < c.M; ++r) { row(c.A); row(c.B); row(c.C); }
    return c;
}

// MatMul witness: X then Y (p^2 each, row-major) from stream 3 of wseed,
// t[i1 p^2 + i2 p + i3] = X[i2][i1] Y[i1][i3], Q = X Y.
std::vector<Fr> make_witness_matmul(const Circuit& c, uint64_t wseed) {
    const size_t P = (size_t)c.mm_p, p2 = P * P;
    std::vector<Fr> z(c.num_vars, Fr::zero());
    z[0] = Fr::one();
    Rng rng(wseed, 3);
    for (size_t i = 1; i <= 2 * p2; ++i) z[i] = rng.field<Fr>();
    const Fr* X = &z[1];
    const Fr* Y = &z[1 + p2];
    for (size_t a = 0; a < P; ++a)
        for (size_t b = 0; b < P; ++b) {
            Fr q = Fr::zero();
            for (size_t k = 0; k < P; ++k) {
                Fr t = X[a * P + k] * Y[k * P + b];
                z[1 + 3 * p2 + k * p2 + a * P + b] = t;
                q += t;
            }
            z[1 + 2 * p2 + a * P + b] = q;
        }
    return z;
}

Fr row_dot(const Csr& m, size_t rows, size_t r, const std::vector<Fr>& z) {
    size_t b = m.row_ptr[r], e = r + 1 < rows ? m.row_ptr[r + 1] : m.col.size();
    Fr acc = Fr::zero();
    for (size_t j = b; j < e; ++j) acc += z[m.col[j]] * m.val[j];
    return acc;
}

std::vector<Fr> make_witness(const Circuit& c, uint64_t wseed) {
    if (c.kind == 1) return make_witness_matmul(c, wseed);
    std::vector<Fr> z(c.num_vars);
    z[0] = Fr::one();
    Rng rng(wseed, 3);
    for (int i = 1; i <= NUM_INPUTS; ++i) z[i] = rng.field<Fr>();
    for (size_t g = 0; g < c.num_gates; ++g) {
        z[1 + NUM_INPUTS + g] = row_dot(c.A, c.M, g, z) * row_dot(c.B, c.M, g, z);
    }
    return z;
}

struct Toxic { Fr tau, alpha, beta, gamma, delta; };
Toxic make_toxic(uint64_t seed) {
    Rng rng(seed, 4);
    Toxic t;
    t.tau = rng.field<Fr>(); t.alpha = rng.field<Fr>(); t.beta = rng.field<Fr>();
    t.gamma = rng.field<Fr>(); t.delta = rng.field<Fr>();
    return t;
}

Fr pow_u64(Fr b, uint64_t e) { uint64_t w[4] = {e, 0, 0, 0}; return b.pow(w, 1); }

// QAP polynomials at tau: At, Bt, Ct for every variable; Z(tau).
void qap_at_tau(const Circuit& c, const Fr& tau, std::vector<Fr>& At, std::vector<Fr>& Bt, std::vector<Fr>& Ct, Fr& Zt) {
    Fr w; orc_fr_root_of_unity(c.log_M, 0, w.v);
    Zt = pow_u64(tau, c.M) - Fr::one();
    Fr Minv = Fr::from_u64(c.M).inverse();
    std::vector<Fr> L(c.M), den(c.M);
    Fr wj = Fr::one();
    for (size_t j = 0; j < c.M; ++j) { L[j] = wj; den[j] = tau - wj; wj *= w; }
    // batch-invert den
    std::vector<Fr> pref(c.M);
    Fr acc = Fr::one();
    for (size_t j = 0; j < c.M; ++j) { pref[j] = acc; acc *= den[j]; }
    Fr inv = acc.inverse();
    for (size_t j = c.M; j-- > 0;) { Fr dj = inv * pref[j]; inv *= den[j]; den[j] = dj; }
    Fr scale = Zt * Minv;
    for (size_t j = 0; j < c.M; ++j) L[j] = L[j] * den[j] * scale;
    At.assign(c.num_vars, Fr::zero()); Bt.assign(c.num_vars, Fr::zero()); Ct.assign(c.num_vars, Fr::zero());
    auto accum = [&](const Csr& m, std::vector<Fr>& out) {
        for (size_t r = 0; r < c.M; ++r) {
            size_t b = m.row_ptr[r], e = r + 1 < c.M ? m.row_ptr[r + 1] : m.col.size();
            for (size_t j = b; j < e; ++j) out[m.col[j]] += m.val[j] * L[r];
        }
    };
    accum(c.A, At); accum(c.B, Bt); accum(c.C, Ct);
}

struct KeyScalars {
    Toxic t;
    Fr Zt;
    std::vector<Fr> At, Bt, Ct, K, IC, H;
};
KeyScalars key_scalars(const Circuit& c, uint64_t seed) {
    KeyScalars k;
    k.t = make_toxic(seed);
    qap_at_tau(c, k.t.tau, k.At, k.Bt, k.Ct, k.Zt);
    Fr dinv = k.t.delta.inverse(), ginv = k.t.gamma.inverse();
    size_t l = c.num_public;
    k.K.resize(c.num_vars - l - 1);
    k.IC.resize(l + 1);
    for (size_t i = 0; i < c.num_vars; ++i) {
        Fr v = k.t.beta * k.At[i] + k.t.alpha * k.Bt[i] + k.Ct[i];
        if (i <= l) k.IC[i] = v * ginv; else k.K[i - l - 1] = v * dinv;
    }
    k.H.resize(c.M - 1);
    Fr tp = k.Zt * dinv;
    for (size_t j = 0; j + 1 < c.M; ++j) { k.H[j] = tp; tp *= k.t.tau; }
    return k;
}

template <class G, class F>
void points_from_scalars(const std::vector<Fr>& s, const FixedBase<G>& fb, std::vector<G>& out) {
    out.resize(s.size());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)s.size(); ++i) { uint64_t c[4]; s[i].to_canonical(c); out[i] = fb.mul(c); }
    batch_to_affine<G, F>(out);
}

void put_g1(const G1& P, uint64_t* out) {   // affine 64 B, infinity = zeros
    G1 A = P.to_affine();
    if (A.is_zero()) std::memset(out, 0, 64);
    else { std::memcpy(out, A.X.v, 32); std::memcpy(out + 4, A.Y.v, 32); }
}
void put_g2(const G2& P, uint64_t* out) {
    G2 A = P.to_affine();
    if (A.is_zero()) { std::memset(out, 0, 128); return; }
    std::memcpy(out, A.X.c0.v, 32); std::memcpy(out + 4, A.X.c1.v, 32);
    std::memcpy(out + 8, A.Y.c0.v, 32); std::memcpy(out + 12, A.Y.c1.v, 32);
}
G1 get_g1(const uint64_t* a) {
    Fq x = Fq::from_raw(a), y = Fq::from_raw(a + 4);
    if (x.is_zero() && y.is_zero()) return G1::zero();
    return {x, y, Fq::one()};
}
G2 get_g2(const uint64_t* a) {
    Fq2 x{Fq::from_raw(a), Fq::from_raw(a + 4)}, y{Fq::from_raw(a + 8), Fq::from_raw(a + 12)};
    if (x.is_zero() && y.is_zero()) return G2::zero();
    return {x, y, Fq2::one()};
}
G1 mul1(const Fr& k, const G1& P) { uint64_t c[4]; k.to_canonical(c); return P.scalar_mul(c); }
G2 mul2(const Fr& k, const G2& P) { uint64_t c[4]; k.to_canonical(c); return P.scalar_mul(c); }

// H coefficients (M entries; h[M-1] = 0 for a satisfying witness).
std::vector<Fr> compute_h(const Circuit& c, const std::vector<Fr>& z) {
    size_t M = c.M;
    std::vector<Fr> a(M), b(M), cc(M);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)M; ++r) {
        a[r] = row_dot(c.A, M, r, z); b[r] = row_dot(c.B, M, r, z); cc[r] = row_dot(c.C, M, r, z);
    }
    Fr w, wi; orc_fr_root_of_unity(c.log_M, 0, w.v); orc_fr_root_of_unity(c.log_M, 1, wi.v);
    std::vector<Fr> tw(M / 2 > 0 ? M / 2 : 1), twi(M / 2 > 0 ? M / 2 : 1);
    tw[0] = twi[0] = Fr::one();
    for (size_t i = 1; i < M / 2; ++i) { tw[i] = tw[i - 1] * w; twi[i] = twi[i - 1] * wi; }
    Fr Minv = Fr::from_u64(M).inverse();
    Fr g = Fr::from_u64(5), gi = g.inverse();
    auto intt = [&](std::vector<Fr>& v) { orc_ntt(M, c.log_M, v[0].v, twi[0].v, Minv.v); };
    auto ntt = [&](std::vector<Fr>& v) { orc_ntt(M, c.log_M, v[0].v, tw[0].v, nullptr); };
    auto shift = [&](std::vector<Fr>& v, const Fr& s) { Fr p = Fr::one(); for (size_t j = 0; j < M; ++j) { v[j] *= p; p *= s; } };
    intt(a); intt(b); intt(cc);
    shift(a, g); shift(b, g); shift(cc, g);
    ntt(a); ntt(b); ntt(cc);
    Fr zinv = (pow_u64(g, M) - Fr::one()).inverse();
    for (size_t j = 0; j < M; ++j) a[j] = (a[j] * b[j] - cc[j]) * zinv;
    intt(a);
    shift(a, gi);
    return a;
}

struct Oracle16 {
    Circuit c;
    KeyScalars ks;
    bool have_points = false;
    bool have_vk = false;
    std::vector<G1> A1, B1, L1, H1, IC1;
    std::vector<G2> B2;
    G1 alpha1, beta1, delta1;
    G2 beta2, gamma2, delta2;
};

// Verifying-key points only (IC and the six toxic-waste points): what the
// pairing check needs, without the O(m) proving-key arrays.
void ensure_vk(Oracle16& o) {
    if (o.have_vk) return;
    static FixedBase<G1> fb1(g1_generator());
    points_from_scalars<G1, Fq>(o.ks.IC, fb1, o.IC1);
    o.alpha1 = mul1(o.ks.t.alpha, g1_generator()).to_affine();
    o.beta1 = mul1(o.ks.t.beta, g1_generator()).to_affine();
    o.delta1 = mul1(o.ks.t.delta, g1_generator()).to_affine();
    o.beta2 = mul2(o.ks.t.beta, g2_generator()).to_affine();
    o.gamma2 = mul2(o.ks.t.gamma, g2_generator()).to_affine();
    o.delta2 = mul2(o.ks.t.delta, g2_generator()).to_affine();
    o.have_vk = true;
}

void ensure_points(Oracle16& o) {
    if (o.have_points) return;
    ensure_vk(o);
    static FixedBase<G1> fb1(g1_generator());
    static FixedBase<G2> fb2(g2_generator());
    points_from_scalars<G1, Fq>(o.ks.At, fb1, o.A1);
    points_from_scalars<G1, Fq>(o.ks.Bt, fb1, o.B1);
    points_from_scalars<G2, Fq2>(o.ks.Bt, fb2, o.B2);
    points_from_scalars<G1, Fq>(o.ks.K, fb1, o.L1);
    points_from_scalars<G1, Fq>(o.ks.H, fb1, o.H1);
    o.have_points = true;
}

template <class G>
G msm_pts(const std::vector<Fr>& s, size_t s_off, const std::vector<G>& P, size_t n) {
    return msm_pippenger<G>(s[s_off].v, n, [&P](size_t i) { return P[i]; });
}

// kind 0: multiplication-gate DAG (param = log_M); kind 1: MatMul (param = p).
bool make_any(int kind, uint64_t seed, int param, int num_public, Circuit& out) {
    if (kind == 1) {
        if (param < 2 || param > 255) return false;
        out = make_circuit_matmul(param);
        return out.log_M <= 24;
    }
    if (kind != 0 || param < 4 || param > 24 || num_public < 0 || (size_t)num_public + 1 >= ((size_t)1 << param)) return false;
    out = make_circuit(seed, param, num_public);
    return true;
}

}  // namespace

extern "C" {

// Circuit of either kind: sizes {M, num_vars, num_public, num_gates, nnzA, nnzB, nnzC}.
int orc_circuit_sizes_k(int kind, uint64_t seed, int param, int num_public, uint64_t* out) {
    Circuit c;
    if (!make_any(kind, seed, param, num_public, c)) return -1;
    out[0] = c.M; out[1] = c.num_vars; out[2] = c.num_public; out[3] = c.num_gates;
    out[4] = c.A.col.size(); out[5] = c.B.col.size(); out[6] = c.C.col.size();
    return 0;
}
int orc_circuit_csr_k(int kind, uint64_t seed, int param, int num_public, int which, uint64_t* row_ptr, uint64_t* col, uint64_t* val) {
    Circuit c;
    if (!make_any(kind, seed, param, num_public, c)) return -1;
    const Csr& m = which == 0 ? c.A : which == 1 ? c.B : c.C;
    std::memcpy(row_ptr, m.row_ptr.data(), 8 * m.row_ptr.size());
    std::memcpy(col, m.col.data(), 8 * m.col.size());
    std::memcpy(val, m.val.data(), 32 * m.val.size());
    return 0;
}
int orc_circuit_witness_k(int kind, uint64_t seed, int param, int num_public, uint64_t wseed, uint64_t* z) {
    Circuit c;
    if (!make_any(kind, seed, param, num_public, c)) return -1;
    std::vector<Fr> w = make_witness(c, wseed);
    std::memcpy(z, w.data(), 32 * w.size());
    return 0;
}
int orc_r1cs_satisfied_k(int kind, uint64_t seed, int param, int num_public, const uint64_t* z) {
    Circuit c;
    if (!make_any(kind, seed, param, num_public, c)) return -1;
    std::vector<Fr> zz(c.num_vars);
    std::memcpy(zz.data(), z, 32 * c.num_vars);
    for (size_t r = 0; r < c.M; ++r)
        if (row_dot(c.A, c.M, r, zz) * row_dot(c.B, c.M, r, zz) != row_dot(c.C, c.M, r, zz)) return 0;
    return 1;
}
void* orc_g16_new_k(int kind, uint64_t seed, int param, int num_public) {
    Circuit c;
    if (!make_any(kind, seed, param, num_public, c)) return nullptr;
    Oracle16* o = new Oracle16();
    o->c = std::move(c);
    o->ks = key_scalars(o->c, seed);
    return o;
}


// Sizes of the synthetic circuit: out = {M, num_vars, num_public, num_gates, nnzA, nnzB, nnzC}.
int orc_circuit_sizes(uint64_t seed, int log_M, int num_public, uint64_t* out) {
    if (log_M < 4 || log_M > 24 || num_public < 0 || (size_t)num_public + 1 >= ((size_t)1 << log_M)) return -1;
    Circuit c = make_circuit(seed, log_M, num_public);
    out[0] = c.M; out[1] = c.num_vars; out[2] = c.num_public; out[3] = c.num_gates;
    out[4] = c.A.col.size(); out[5] = c.B.col.size(); out[6] = c.C.col.size();
    return 0;
}
// CSR of matrix `which` (0 A, 1 B, 2 C): row_ptr[M] (no sentinel), col[nnz], val[nnz] Fr.
int orc_circuit_csr(uint64_t seed, int log_M, int num_public, int which, uint64_t* row_ptr, uint64_t* col, uint64_t* val) {
    Circuit c = make_circuit(seed, log_M, num_public);
    const Csr& m = which == 0 ? c.A : which == 1 ? c.B : c.C;
    std::memcpy(row_ptr, m.row_ptr.data(), 8 * m.row_ptr.size());
    std::memcpy(col, m.col.data(), 8 * m.col.size());
    std::memcpy(val, m.val.data(), 32 * m.val.size());
    return 0;
}
int orc_circuit_witness(uint64_t seed, int log_M, int num_public, uint64_t wseed, uint64_t* z) {
    Circuit c = make_circuit(seed, log_M, num_public);
    std::vector<Fr> w = make_witness(c, wseed);
    std::memcpy(z, w.data(), 32 * w.size());
    return 0;
}
// Host R1CS check: (A z) o (B z) == C z on every row.
int orc_r1cs_satisfied(uint64_t seed, int log_M, int num_public, const uint64_t* z) {
    Circuit c = make_circuit(seed, log_M, num_public);
    std::vector<Fr> zz(c.num_vars);
    std::memcpy(zz.data(), z, 32 * c.num_vars);
    for (size_t r = 0; r < c.M; ++r)
        if (row_dot(c.A, c.M, r, zz) * row_dot(c.B, c.M, r, zz) != row_dot(c.C, c.M, r, zz)) return 0;
    return 1;
}

void* orc_g16_new(uint64_t seed, int log_M, int num_public) {
    Oracle16* o = new Oracle16();
    o->c = make_circuit(seed, log_M, num_public);
    o->ks = key_scalars(o->c, seed);
    return o;
}
void orc_g16_free(void* h) { delete static_cast<Oracle16*>(h); }

// Key scalars (trapdoor view): which 0 At, 1 Bt, 2 Ct, 3 K (L query), 4 IC, 5 H query,
// 6 toxic {tau, alpha, beta, gamma, delta}. Returns the element count.
size_t orc_g16_key_scalars(void* h, int which, uint64_t* out) {
    Oracle16* o = static_cast<Oracle16*>(h);
    const std::vector<Fr>* v = nullptr;
    switch (which) {
        case 0: v = &o->ks.At; break; case 1: v = &o->ks.Bt; break; case 2: v = &o->ks.Ct; break;
        case 3: v = &o->ks.K; break; case 4: v = &o->ks.IC; break; case 5: v = &o->ks.H; break;
        case 6: {
            const Toxic& t = o->ks.t;
            const Fr* xs[5] = {&t.tau, &t.alpha, &t.beta, &t.gamma, &t.delta};
            if (out) for (int i = 0; i < 5; ++i) std::memcpy(out + 4 * i, xs[i]->v, 32);
            return 5;
        }
        default: return 0;
    }
    if (out) std::memcpy(out, v->data(), 32 * v->size());
    return v->size();
}

// Key points (affine): which 0 A1, 1 B1, 2 B2 (128 B), 3 L1, 4 H1, 5 IC1,
// 6 {alpha1, beta1, delta1} (3 x 64 B), 7 {beta2, gamma2, delta2} (3 x 128 B).
size_t orc_g16_key_points(void* h, int which, uint64_t* out) {
    Oracle16* o = static_cast<Oracle16*>(h);
    ensure_points(*o);
    auto dump1 = [&](const std::vector<G1>& v) { if (out) for (size_t i = 0; i < v.size(); ++i) put_g1(v[i], out + 8 * i); return v.size(); };
    switch (which) {
        case 0: return dump1(o->A1);
        case 1: return dump1(o->B1);
        case 2: if (out) for (size_t i = 0; i < o->B2.size(); ++i) put_g2(o->B2[i], out + 16 * i); return o->B2.size();
        case 3: return dump1(o->L1);
        case 4: return dump1(o->H1);
        case 5: return dump1(o->IC1);
        case 6: if (out) { put_g1(o->alpha1, out); put_g1(o->beta1, out + 8); put_g1(o->delta1, out + 16); } return 3;
        case 7: if (out) { put_g2(o->beta2, out); put_g2(o->gamma2, out + 16); put_g2(o->delta2, out + 32); } return 3;
        default: return 0;
    }
}

// H coefficients for witness z (M entries).
int orc_g16_h(void* h, const uint64_t* z, uint64_t* out) {
    Oracle16* o = static_cast<Oracle16*>(h);
    std::vector<Fr> zz(o->c.num_vars);
    std::memcpy(zz.data(), z, 32 * zz.size());
    std::vector<Fr> hh = compute_h(o->c, zz);
    std::memcpy(out, hh.data(), 32 * hh.size());
    return 0;
}

// r, s for a witness seed (stream 5).
void orc_g16_rs(uint64_t wseed, uint64_t* out) {
    Rng rng(wseed, 5);
    Fr r = rng.field<Fr>(), s = rng.field<Fr>();
    std::memcpy(out, r.v, 32); std::memcpy(out + 4, s.v, 32);
}

// Trapdoor closed-form proof: proof = {A: 64 B, B: 128 B, C: 64 B} affine.
int orc_g16_prove_trapdoor(void* h, const uint64_t* z, const uint64_t* rs, uint64_t* proof) {
    Oracle16* o = static_cast<Oracle16*>(h);
    const KeyScalars& k = o->ks;
    Fr r = Fr::from_raw(rs), s = Fr::from_raw(rs + 4);
    Fr Az = Fr::zero(), Bz = Fr::zero(), Cz = Fr::zero(), Kz = Fr::zero();
    for (size_t i = 0; i < o->c.num_vars; ++i) {
        Fr zi = Fr::from_raw(z + 4 * i);
        Az += zi * k.At[i]; Bz += zi * k.Bt[i]; Cz += zi * k.Ct[i];
        if (i > o->c.num_public) Kz += zi * k.K[i - o->c.num_public - 1];
    }
    Fr a = k.t.alpha + Az + r * k.t.delta;
    Fr b = k.t.beta + Bz + s * k.t.delta;
    Fr hz = (Az * Bz - Cz) * k.t.delta.inverse();   // h(tau) Z(tau) / delta
    Fr c = Kz + hz + s * a + r * b - r * s * k.t.delta;
    put_g1(mul1(a, g1_generator()), proof);
    put_g2(mul2(b, g2_generator()), proof + 8);
    put_g1(mul1(c, g1_generator()), proof + 24);
    return 0;
}

// Honest prover on the CPU: SpMV, NTT pipeline for H, CPU Pippenger MSMs, phase-6 assembly.
int orc_g16_prove(void* h, const uint64_t* z, const uint64_t* rs, uint64_t* proof) {
    Oracle16* o = static_cast<Oracle16*>(h);
    ensure_points(*o);
    const Circuit& c = o->c;
    std::vector<Fr> zz(c.num_vars);
    std::memcpy(zz.data(), z, 32 * zz.size());
    std::vector<Fr> hh = compute_h(c, zz);
    Fr r = Fr::from_raw(rs), s = Fr::from_raw(rs + 4);
    size_t l = c.num_public;
    G1 sA = msm_pts<G1>(zz, 0, o->A1, c.num_vars);
    G1 sB1 = msm_pts<G1>(zz, 0, o->B1, c.num_vars);
    G2 sB2 = msm_pts<G2>(zz, 0, o->B2, c.num_vars);
    G1 sL = msm_pts<G1>(zz, l + 1, o->L1, c.num_vars - l - 1);
    G1 sH = msm_pts<G1>(hh, 0, o->H1, c.M - 1);
    G1 A = o->alpha1 + sA + mul1(r, o->delta1);
    G2 B = o->beta2 + sB2 + mul2(s, o->delta2);
    G1 B1 = o->beta1 + sB1 + mul1(s, o->delta1);
    G1 C = sL + sH + mul1(s, A) + mul1(r, B1) + mul1(r * s, o->delta1).neg();
    put_g1(A, proof); put_g2(B, proof + 8); put_g1(C, proof + 24);
    return 0;
}

// Pairing verifier; returns 1 if the proof is accepted.
int orc_g16_verify(void* h, const uint64_t* public_z /* num_public+1 values, z[0]=1 */, const uint64_t* proof) {
    Oracle16* o = static_cast<Oracle16*>(h);
    ensure_vk(*o);
    G1 vkx = G1::zero();
    for (size_t i = 0; i <= o->c.num_public; ++i) vkx = vkx + mul1(Fr::from_raw(public_z + 4 * i), o->IC1[i]);
    uint64_t ps[4 * 8], qs[4 * 16];
    G1 A = get_g1(proof);
    G2 B = get_g2(proof + 8);
    G1 C = get_g1(proof + 24);
    put_g1(A.neg(), ps); put_g2(B, qs);
    put_g1(o->alpha1, ps + 8); put_g2(o->beta2, qs + 16);
    put_g1(vkx, ps + 16); put_g2(o->gamma2, qs + 32);
    put_g1(C, ps + 24); put_g2(o->delta2, qs + 48);
    return orc_pairing_product_is_one(ps, qs, 4);
}

}  // extern "C"
