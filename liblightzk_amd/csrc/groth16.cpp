// Groth16 over alt_bn128 on one GPU: synthetic circuit, setup from seeded
// toxic waste, and the prover.
//
// Keeps the reference's protocol shape (include/protocol/groth16.cuh:18-223:
// construct, Setup, Prove with phases SpMV -> NTTs -> MSMs -> host assembly)
// but proves a real statement, which the reference does not (random keys and
// witness, no verifier: groth16.cuh:84-132):
//   circuit  multiplication-gate DAG: gate g defines z_v = (A_g . z)(B_g . z),
//            v = 65 + g, A/B rows hold 1-3 entries with columns uniform over
//            earlier variables and Fr coefficients; the last num_public+1
//            rows hold z_i in A only (input-consistency rows); M = 2^log_m
//            rows; z_0 = 1, z_1..z_64 seeded (first num_public are public).
//   setup    QAP at tau: L_j(tau) = (tau^M - 1) w^j / (M (tau - w^j)),
//            A_i(tau) = sum_j A[j][i] L_j(tau) (likewise B, C) on the host;
//            key points k*G by the GPU fixed-base comb and kept in HBM:
//            A_i G1, B_i G1, B_i G2, (beta A_i + alpha B_i + C_i)/delta G1
//            (i > num_public), tau^j Z(tau)/delta G1 (j < M-1).
//   prove    SpMV (A z, B z, C z) -> 3 iNTT -> 3 coset NTT ->
//            (A B - C)/(g^M - 1) -> coset iNTT = h -> MSMs over z and h ->
//            phase-6 assembly on the host (groth16.cuh:210-215):
//            A = alpha1 + sum z_i A_i + r delta1
//            B = beta2 + sum z_i B2_i + s delta2,  B1 = beta1 + sum z_i B1_i + s delta1
//            C = sum_{i>l} z_i L_i + sum_j h_j H_j + s A + r B1 - r s delta1
//              = [sum_{i>l} z_i L_i + sum_j h_j H_j + sum_i (r z_i) B1_i] + s A + r beta1
//            The MSMs skip the key's points at infinity (compacted A1, B1,
//            B2) and the bracket is one merged MSM; see g16.hpp.
// Streams: 2 circuit structure, 3 witness inputs, 4 toxic waste
// (tau, alpha, beta, gamma, delta), 5 prover randomness (r, s).
#include <omp.h>
#include <algorithm>
#include <thread>
#include <vector>
#include "common.hpp"
#include "host_ec.hpp"
#include "msm.hpp"
#include "synth.hpp"
#include "g16.hpp"

using namespace hzk;
using host::Fr;
using host::Fq;
using host::G1x;
using host::G2x;

namespace {

const int NUM_INPUTS = 64;

Fr fr_u64(uint64_t v) {
    Fr t{{v, 0, 0, 0}};
    Fr r2;
    std::memcpy(r2.v, host::FR.r2, 32);
    return t * r2;
}
Fr fr_draw(synth::Xoshiro& g) {
    Fr r;
    synth::draw(g, synth::FR_MOD, r.v);
    return r;
}
Fr fr_root(int k) {
    static const uint64_t t[4] = {0x9b9709143e1f593fULL, 0x181585d2833e8487ULL, 0x131a029b85045b68ULL, 0x000000030644e72eULL};
    Fr w = fr_u64(5).pow(t);
    for (int i = 28; i > k; --i) w = w.sqr();
    return w;
}
Fr fr_pow_u64(Fr b, uint64_t e) {
    Fr acc = Fr::one();
    while (e) {
        if (e & 1) acc = acc * b;
        b = b.sqr();
        e >>= 1;
    }
    return acc;
}

}  // namespace

namespace {

void build_circuit(hzk_g16* g, uint64_t seed) {
    synth::Xoshiro rng(seed, 2);
    HostCsr &A = g->m[0], &B = g->m[1], &C = g->m[2];
    for (size_t gate = 0; gate < g->num_gates; ++gate) {
        uint64_t v = 1 + NUM_INPUTS + gate;
        for (HostCsr* mm : {&A, &B}) {
            mm->row_ptr.push_back(mm->col.size());
            int k = (int)(rng.next() % 3) + 1;
            for (int t = 0; t < k; ++t) {
                mm->col.push_back(rng.next() % v);
                mm->val.push_back(fr_draw(rng));
            }
        }
        C.row_ptr.push_back(C.col.size());
        C.col.push_back(v);
        C.val.push_back(Fr::one());
    }
    for (size_t i = 0; i <= g->num_public; ++i) {
        A.row_ptr.push_back(A.col.size());
        A.col.push_back(i);
        A.val.push_back(Fr::one());
        B.row_ptr.push_back(B.col.size());
        C.row_ptr.push_back(C.col.size());
    }
}

// The reference's MatMul R1CS for p x p matrices (xpu-spmat.cuh:33-65):
// variables z_0 = 1, X at 1 + a p + b, Y at 1 + p^2 + a p + b, the product
// matrix Q at 1 + 2 p^2 + a p + b, and p^3 partial products t_i at
// 1 + 3 p^2 + i. Row i < p^3 (i = i1 p^2 + i2 p + i3) is
// X[i2][i1] * Y[i1][i3] = t_i; row p^3 + i (i = i2 p + i3) is
// (sum_k t_{k p^2 + i}) * z_0 = Q[i2][i3]; rows up to M = next_pow2(p^3 + p^2)
// are empty. All coefficients are one. The public inputs are X, Y and Q
// (3 p^2 of them, the reference's k).
void build_circuit_matmul(hzk_g16* g) {
    const size_t p = (size_t)g->mm_p, p2 = p * p, p3 = p2 * p;
    HostCsr &A = g->m[0], &B = g->m[1], &C = g->m[2];
    for (HostCsr* mm : {&A, &B, &C}) {
        mm->row_ptr.reserve(g->M);
        mm->col.reserve(2 * p3);
        mm->val.reserve(2 * p3);
    }
    auto put = [](HostCsr& mm, uint64_t col) {
        mm.col.push_back(col);
        mm.val.push_back(Fr::one());
    };
    for (size_t i = 0; i < p3; ++i) {
        const size_t i1 = i / p2, i2 = (i % p2) / p, i3 = i % p;
        for (HostCsr* mm : {&A, &B, &C}) mm->row_ptr.push_back(mm->col.size());
        put(A, 1 + i2 * p + i1);
        put(B, 1 + p2 + i1 * p + i3);
        put(C, 1 + 3 * p2 + i);
    }
    for (size_t i = 0; i < p2; ++i) {
        for (HostCsr* mm : {&A, &B, &C}) mm->row_ptr.push_back(mm->col.size());
        for (size_t k = 0; k < p; ++k) put(A, 1 + 3 * p2 + k * p2 + i);
        put(B, 0);
        put(C, 1 + 2 * p2 + i);
    }
    for (size_t r = p3 + p2; r < g->M; ++r)
        for (HostCsr* mm : {&A, &B, &C}) mm->row_ptr.push_back(mm->col.size());
}

Fr row_dot(const HostCsr& m, size_t rows, size_t r, const Fr* z) {
    size_t b = m.row_ptr[r], e = r + 1 < rows ? m.row_ptr[r + 1] : m.col.size();
    Fr acc = Fr::zero();
    for (size_t j = b; j < e; ++j) acc = acc + z[m.col[j]] * m.val[j];
    return acc;
}

template <class T>
int dev_upload(hzk_ctx* ctx, hzk_g16* g, const T* src, size_t bytes, void** out) {
    void* d = nullptr;
    HZK_HIP(ctx, hipMalloc(&d, bytes ? bytes : 1));
    g->own(d, bytes);
    if (bytes) HZK_HIP(ctx, hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
    *out = d;
    return 0;
}

int dev_alloc(hzk_ctx* ctx, hzk_g16* g, size_t bytes, void** out) {
    void* d = nullptr;
    HZK_HIP(ctx, hipMalloc(&d, bytes ? bytes : 1));
    g->own(d, bytes);
    *out = d;
    return 0;
}

// k_i * G for host scalars k into a new device array of affine points
template <bool G2>
int key_points(hzk_ctx* ctx, hzk_g16* g, const std::vector<Fr>& k, void** out) {
    size_t n = k.size();
    int rc = dev_alloc(ctx, g, n * (G2 ? 128 : 64), out);
    if (rc) return rc;
    if (n == 0) return 0;
    void* ds = nullptr;
    HZK_HIP(ctx, hipMalloc(&ds, n * 32));
    HZK_HIP(ctx, hipMemcpy(ds, k.data(), n * 32, hipMemcpyHostToDevice));
    rc = G2 ? hzk_fixbase_g2(ctx, ds, n, *out) : hzk_fixbase_g1(ctx, ds, n, *out);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ds);
    return rc;
}

// Window size of the key's level tables: minimises bucket entries
// n * ceil(255 / c) plus the reduction of its bucket sets, 3 * 2^(c-1) each.
int level_window(size_t n, uint32_t levels) {
    int best = 16;
    double bc = -1;
    for (int c = MSM_MIN_C; c <= 16; ++c) {
        const uint32_t W = (255 + c - 1) / c, sets = (W + levels - 1) / levels;
        double cost = (double)n * W + 3.0 * sets * (double)(1u << (c - 1));
        if (bc < 0 || cost < bc) { bc = cost; best = c; }
    }
    return best;
}

// Level tables of the three MSM base arrays (compacted A1, the merged C1,
// compacted B2) with `levels` levels each (point t*n + i = 2^(c t) P_i): a
// proof's MSMs then reduce ceil(W / levels) bucket sets each instead of W.
// levels <= 1 frees them (plain arrays).
int build_levels(hzk_ctx* ctx, hzk_g16* g, uint32_t levels) {
    for (void** p : {&g->A1L, &g->C1L, &g->B2L}) {
        if (*p) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(*p);
            g->disown(*p);
            *p = nullptr;
        }
    }
    g->lv = 1;
    g->lc = 0;
    if (levels <= 1) return 0;
    const int c = level_window(g->num_vars, levels);
    const uint32_t lv = std::min<uint32_t>(levels, (uint32_t)((255 + c - 1) / c));
    struct T { void* src; size_t n; int group; void** dst; };
    const T tabs[3] = {{g->A1c, g->nA, 1, &g->A1L}, {g->C1, g->nC, 1, &g->C1L}, {g->B2c, g->nB, 2, &g->B2L}};
    for (const T& t : tabs) {
        int rc = dev_alloc(ctx, g, t.n * lv * (t.group == 1 ? 64 : 128), t.dst);
        if (!rc) rc = hzk_bases_levels(ctx, t.group, c, (int)lv, t.src, t.n, *t.dst);
        if (!rc) rc = t.group == 1 ? msm_convert_g1(ctx, *t.dst, *t.dst, t.n * lv) : msm_convert_g2(ctx, *t.dst, *t.dst, t.n * lv);
        if (rc) return rc;
    }
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    g->lc = c;
    g->lv = lv;
    return 0;
}

int setup(hzk_ctx* ctx, hzk_g16* g, uint64_t seed) {
    if (g->kind == 1) build_circuit_matmul(g);
    else build_circuit(g, seed);
    // device CSR
    for (int i = 0; i < 3; ++i) {
        int rc = dev_upload(ctx, g, g->m[i].val.data(), g->m[i].val.size() * 32, &g->d_val[i]);
        if (!rc) rc = dev_upload(ctx, g, g->m[i].col.data(), g->m[i].col.size() * 8, (void**)&g->d_col[i]);
        if (!rc) rc = dev_upload(ctx, g, g->m[i].row_ptr.data(), g->m[i].row_ptr.size() * 8, (void**)&g->d_rp[i]);
        if (rc) return rc;
    }
    // toxic waste
    synth::Xoshiro tw(seed, 4);
    Fr tau = fr_draw(tw), alpha = fr_draw(tw), beta = fr_draw(tw), gamma = fr_draw(tw), delta = fr_draw(tw);
    const size_t M = g->M, N1 = g->num_vars, l = g->num_public;
    // Lagrange basis at tau
    Fr w = fr_root(g->log_m);
    Fr Zt = fr_pow_u64(tau, M) - Fr::one();
    Fr scale = Zt * fr_u64(M).inverse();
    std::vector<Fr> L(M), den(M), pre(M);
    Fr wj = Fr::one();
    for (size_t j = 0; j < M; ++j) { L[j] = wj; den[j] = tau - wj; wj = wj * w; }
    Fr acc = Fr::one();
    for (size_t j = 0; j < M; ++j) { pre[j] = acc; acc = acc * den[j]; }
    HZK_CHECK(ctx, !acc.is_zero(), HZK_ERR_INVALID, "setup: tau lies in the evaluation domain");
    Fr inv = acc.inverse();
    for (size_t j = M; j-- > 0;) { Fr dj = inv * pre[j]; inv = inv * den[j]; L[j] = L[j] * dj * scale; }
    std::vector<Fr> At(N1, Fr::zero()), Bt(N1, Fr::zero()), Ct(N1, Fr::zero());
    std::vector<Fr>* outs[3] = {&At, &Bt, &Ct};
    for (int i = 0; i < 3; ++i) {
        const HostCsr& mm = g->m[i];
        for (size_t r = 0; r < M; ++r) {
            size_t b = mm.row_ptr[r], e = r + 1 < M ? mm.row_ptr[r + 1] : mm.col.size();
            for (size_t j = b; j < e; ++j) (*outs[i])[mm.col[j]] = (*outs[i])[mm.col[j]] + mm.val[j] * L[r];
        }
    }
    Fr dinv = delta.inverse(), ginv = gamma.inverse();
    std::vector<Fr> K(N1 - l - 1), IC(l + 1), H(M - 1);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)N1; ++i) {
        Fr v = beta * At[i] + alpha * Bt[i] + Ct[i];
        if ((size_t)i <= l) IC[i] = v * ginv;
        else K[i - l - 1] = v * dinv;
    }
    Fr tp = Zt * dinv;
    for (size_t j = 0; j + 1 < M; ++j) { H[j] = tp; tp = tp * tau; }
    int rc = key_points<false>(ctx, g, At, &g->A1);
    if (!rc) rc = key_points<false>(ctx, g, Bt, &g->B1);
    if (!rc) rc = key_points<true>(ctx, g, Bt, &g->B2);
    if (!rc) rc = key_points<false>(ctx, g, K, &g->L1);
    if (!rc) rc = key_points<false>(ctx, g, H, &g->H1);
    if (rc) return rc;
    auto canon = [](const Fr& x, uint64_t c[4]) { x.canonical(c); };
    uint64_t c[4];
    G1x G = host::g1_generator();
    G2x G2 = host::g2_generator();
    canon(alpha, c); g->alpha1 = G.mul(c);
    canon(beta, c); g->beta1 = G.mul(c); g->beta2 = G2.mul(c);
    canon(delta, c); g->delta1 = G.mul(c); g->delta2 = G2.mul(c);
    canon(gamma, c); g->gamma2 = G2.mul(c);
    g->IC.resize(l + 1);
#pragma omp parallel for schedule(static)
    for (long i = 0; i <= (long)l; ++i) {
        uint64_t ci[4];
        IC[i].canonical(ci);
        g->IC[i] = G.mul(ci);
    }
    rc = g16_sparse_key(ctx, g);
    if (rc) return rc;
    return g16_prover_buffers(ctx, g);
}

}  // namespace

// dst[p] = src[rev_k(p)], p < n, for 64-byte affine G1 points (n <= 2^k - 1,
// so rev_k(p) < 2^k - 1 as well)
__global__ void __launch_bounds__(256) k_gather_bitrev64(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n, int k) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // one 16-byte word
    if (q >= 4 * n) return;
    const size_t p = q >> 2;
    const size_t r = __brev((uint32_t)p) >> (32 - k);
    dst[q] = src[4 * r + (q & 3)];
}

// Compacted A and B query arrays: the variables whose A1 (B1) point is not
// the point at infinity (B_i(tau) = 0 makes B1_i and B2_i infinity alike),
// found on the host from the key arrays, then gathered on the device; and
// the merged base array C1 = L1 | H1 (bit-reversed order, matching the
// quotient h that the H pipeline leaves bit-reversed) | B1c of the C MSM.
int hzk::g16_sparse_key(hzk_ctx* ctx, hzk_g16* g) {
    const size_t N1 = g->num_vars;
    std::vector<uint64_t> pts(N1 * 8);
    std::vector<uint32_t> ia, ib;
    for (int which = 0; which < 2; ++which) {
        HZK_HIP(ctx, hipMemcpy(pts.data(), which ? g->B1 : g->A1, N1 * 64, hipMemcpyDeviceToHost));
        std::vector<uint32_t>& idx = which ? ib : ia;
        idx.reserve(N1);
        for (size_t i = 0; i < N1; ++i) {
            const uint64_t* p = &pts[8 * i];
            if (p[0] | p[1] | p[2] | p[3] | p[4] | p[5] | p[6] | p[7]) idx.push_back((uint32_t)i);
        }
        if (idx.empty()) idx.push_back(0);   // an MSM needs one point: infinity contributes nothing
    }
    const size_t nL = N1 - g->num_public - 1, nH = g->M - 1;
    g->nA = ia.size();
    g->nB = ib.size();
    g->nC = nL + nH + g->nB;
    g->vec_b = std::max(g->M, g->nA + g->nB);
    g->vec_c = std::max(g->M, g->nC);
    int rc = dev_upload(ctx, g, ia.data(), ia.size() * 4, (void**)&g->idxA);
    if (!rc) rc = dev_upload(ctx, g, ib.data(), ib.size() * 4, (void**)&g->idxB);
    if (!rc) rc = dev_alloc(ctx, g, g->nA * 64, &g->A1c);
    if (!rc) rc = dev_alloc(ctx, g, g->nB * 128, &g->B2c);
    if (!rc) rc = dev_alloc(ctx, g, g->nC * 64, &g->C1);
    if (!rc) rc = gather_rows(ctx, g->A1, g->idxA, g->nA, 64, g->A1c);
    if (!rc) rc = gather_rows(ctx, g->B2, g->idxB, g->nB, 128, g->B2c);
    // C1 = L1 | H1 bit-reversed | B1 at idxB
    if (!rc) rc = gather_rows(ctx, g->B1, g->idxB, g->nB, 64, (char*)g->C1 + (nL + nH) * 64);
    if (rc) return rc;
    HZK_HIP(ctx, hipMemcpyAsync(g->C1, g->L1, nL * 64, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_gather_bitrev64, dim3(blocks_for(4 * nH, 256)), dim3(256), 0, ctx->stream, (const uint4*)g->H1,
                       (uint4*)((char*)g->C1 + nL * 64), nH, g->log_m);
    HZK_HIP(ctx, hipGetLastError());
    // the G1 MSM arrays in the accumulation's form
    rc = dev_alloc(ctx, g, g->nA * 64, &g->A1cv);
    if (!rc) rc = dev_alloc(ctx, g, g->nC * 64, &g->C1v);
    if (!rc) rc = dev_alloc(ctx, g, g->nB * 128, &g->B2cv);
    if (!rc) rc = msm_convert_g1(ctx, g->A1c, g->A1cv, g->nA);
    if (!rc) rc = msm_convert_g1(ctx, g->C1, g->C1v, g->nC);
    if (!rc) rc = msm_convert_g2(ctx, g->B2c, g->B2cv, g->nB);
    if (rc) return rc;
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HZK_OK;
}

// Buffers of the plain prover (z and the three evaluation vectors) and
// 1/(g^M - 1) for the coset division.
int hzk::g16_prover_buffers(hzk_ctx* ctx, hzk_g16* g) {
    const size_t M = g->M;
    int rc = dev_alloc(ctx, g, g->num_vars * 32, &g->z);
    if (!rc) rc = dev_alloc(ctx, g, M * 32, &g->a);
    if (!rc) rc = dev_alloc(ctx, g, g->vec_b * 32, &g->b);
    if (!rc) rc = dev_alloc(ctx, g, g->vec_c * 32, &g->c);
    if (rc) return rc;
    Fr zi = (fr_pow_u64(fr_u64(5), M) - Fr::one()).inverse();
    g->fb_delta1.build(g->delta1);
    g->fb_beta1.build(g->beta1);
    g->fb_delta2.build(g->delta2);
    return dev_upload(ctx, g, zi.v, 32, &g->zinv);
}

namespace hzk {

int g16_h_batch(hzk_ctx* ctx, G16Req* reqs, size_t n) {
    std::vector<NttJob> jobs;
    {   // the three SpMVs of every request: one launch
        std::vector<hzk_csr> cs(3 * n);
        std::vector<SpmvJob> sj(n);
        for (size_t i = 0; i < n; ++i) {
            hzk_g16* g = reqs[i].g;
            for (int m = 0; m < 3; ++m) {
                cs[3 * i + m] = {g->d_val[m], g->d_col[m], g->d_rp[m], g->M, g->num_vars, g->m[m].col.size()};
                sj[i].m[m] = &cs[3 * i + m];
            }
            sj[i].z = reqs[i].z;
            sj[i].out[0] = reqs[i].a, sj[i].out[1] = reqs[i].b, sj[i].out[2] = reqs[i].c;
        }
        int rc = spmv3_batch(ctx, sj.data(), n);
        if (rc) return rc;
    }
    // evaluations on <w> (natural order) -> coefficients (bit-reversed, times M)
    // -> evaluations on g<w> (natural order; g^rev(p) / M applied on the way
    // in) -> (A B - C) / (g^M - 1) -> coefficients of h, bit-reversed (g^-rev(p)
    // / M applied on the way out): no bit-reverse permutation anywhere
    for (int kind : {4, 5}) {
        jobs.clear();
        for (size_t i = 0; i < n; ++i)
            for (void* v : {reqs[i].a, reqs[i].b, reqs[i].c}) jobs.push_back({v, reqs[i].g->M, kind});
        int rc = ntt_batch(ctx, jobs.data(), jobs.size(), 0);
        if (rc) return rc;
    }
    jobs.clear();
    std::vector<AbcJob> aj(n);
    for (size_t i = 0; i < n; ++i) {
        aj[i] = {reqs[i].a, reqs[i].b, reqs[i].c, reqs[i].g->zinv, reqs[i].g->M};
        jobs.push_back({reqs[i].a, reqs[i].g->M, 6});
    }
    int rc = pointwise_abc_batch(ctx, aj.data(), n);
    if (rc) return rc;
    return ntt_batch(ctx, jobs.data(), jobs.size(), 0);
}

int g16_msm_batch(hzk_ctx* ctx, G16Req* reqs, size_t n, int graph) {
    std::vector<MsmJob> j1, j2;
    std::vector<GatherJob> gj;
    for (size_t i = 0; i < n; ++i) {
        hzk_g16* g = reqs[i].g;
        const size_t N1 = g->num_vars, l = g->num_public, nL = N1 - l - 1, nH = g->M - 1;
        char* r = reqs[i].res;
        // scalars into the buffers the H stage no longer needs (one launch for
        // every request): b = z_A | z_B, c = z_{l+1..} | h | r z_B
        char* b = (char*)reqs[i].b;
        char* c = (char*)reqs[i].c;
        gj.push_back({reqs[i].z, g->idxA, b, g->nA, 32, nullptr});
        gj.push_back({reqs[i].z, g->idxB, b + g->nA * 32, g->nB, 32, nullptr});
        gj.push_back({reqs[i].z, g->idxB, c + (nL + nH) * 32, g->nB, 32, reqs[i].rs});
        gj.push_back({(const char*)reqs[i].z + (l + 1) * 32, nullptr, c, nL, 32, nullptr});
        gj.push_back({reqs[i].a, nullptr, c + nL * 32, nH, 32, nullptr});
        const uint32_t wc = g->lv > 1 ? (uint32_t)g->lc : 0, L = g->lv > 1 ? g->lv : 0;
        j1.push_back({b, g->lv > 1 ? g->A1L : g->A1cv, g->nA, r, wc, L, 1});
        j1.push_back({c, g->lv > 1 ? g->C1L : g->C1v, g->nC, r + 96, wc, L, 1});
        j2.push_back({b + g->nA * 32, g->lv > 1 ? g->B2L : g->B2cv, g->nB, r + 192, wc, L, 1});
    }
    {
        int rc = gather_batch(ctx, gj.data(), gj.size());
        if (rc) return rc;
    }
    // the G2 batch runs on the auxiliary context's stream, concurrently with
    // the G1 batch: the latency-bound tails of one (bucket reduction, window
    // sums, host combine) overlap the throughput-bound accumulation of the other
    hzk_ctx* aux = nullptr;
    int rc = ctx_aux(ctx, &aux);
    if (rc) return rc;
    hipEvent_t ready;
    HZK_HIP(ctx, hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ready, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(aux->stream, ready, 0);
    (void)hipEventDestroy(ready);
    HZK_HIP(ctx, e);
    int rc2 = 0;
    std::thread t2([&] {
        (void)hipSetDevice(aux->device);
        rc2 = msm_batch_g2(aux, j2.data(), j2.size(), 1, graph);
    });
    rc = msm_batch_g1(ctx, j1.data(), j1.size(), 1, graph);
    t2.join();
    if (ctx->timing) timing_merge(ctx, aux);
    if (!rc && rc2) return set_error(ctx, rc2, "%s", hzk_last_error(aux));
    return rc;
}

// Upper bound of one MSM's scratch (digits and sorted entries, histograms,
// key scans, bucket sums, chunk partials, bucket-reduction and window sums)
// for n points with xyzz points of pt_bytes (bucket sums, chunk partials and
// run sums are raw 29-bit accumulator states, 9/8 of that), from the
// built-in plans (the larger of the single-MSM and wide-batch regimes).
static size_t msm_scratch_bound(size_t n, size_t pt_bytes) {
    const size_t raw = pt_bytes / 8 * 9;
    size_t worst = 0;
    for (int regime : {MSM_REGIME_SINGLE, MSM_REGIME_BATCH}) {
        const MsmPlan p = msm_plan_builtin(pt_bytes > 128, n, regime, 160 * 1024);
        size_t W = p.W, nk = W * p.NB, entries = n * W;
        size_t chunks = (entries + p.K - 1) / p.K;
        size_t bred = W * ((p.NB + p.g - 1) / p.g);
        const size_t coarse = p.NB >= 2048 && n >= MSM_TP_BATCH_MIN_N ? entries * 8 : 0;   // two-pass staging
        size_t b = entries * 8 + coarse + (size_t)p.T * nk * 4 + 2 * (nk + 1) * 4 +
                   2 * ((nk + 1 + MSM_SCAN_TILE) / MSM_SCAN_TILE) * 8 + (nk + 1) * raw + chunks * (4 + 2 * raw) +
                   2 * bred * raw + 2 * 64 * W * pt_bytes + 4096;   // run (A, R) and part (P, RS) pairs
        worst = std::max(worst, b);
    }
    return worst;
}

size_t g16_request_footprint(const hzk_g16* g) {
    // z, a, b, c and the scratch of its three MSMs (A, merged C, B2) with 25% slack
    size_t bufs = (g->num_vars + g->M + g->vec_b + g->vec_c) * 32;
    size_t msm = msm_scratch_bound(g->nA, 128) + msm_scratch_bound(g->nC, 128) + msm_scratch_bound(g->nB, 256);
    return bufs + msm + msm / 4;
}

G16Fixed g16_fixed_terms(const hzk_g16* g, const void* rs) {
    Fr r = Fr::load(rs), s = Fr::load((const char*)rs + 32);
    uint64_t rc[4], sc[4];
    r.canonical(rc);
    s.canonical(sc);
    return {g->fb_delta1.mul(rc), g->fb_beta1.mul(rc), g->fb_delta2.mul(sc)};
}

// Phase 6 on the host; proof = A (64 B) | B (128 B) | C (64 B), affine.
// With the merged C MSM (sum L + sum H + sum (r z_i) B1_i):
//   A = alpha1 + sum A + r delta1, B = beta2 + sum B2 + s delta2,
//   C = merged + s A + r beta1
// (= sum L + sum H + s A + r B1 - r s delta1 with B1 = beta1 + sum B1 + s delta1,
// groth16.cuh:210-215). fx holds r delta1, s delta2 and r beta1.
void g16_assemble(const hzk_g16* g, const char* res, const void* rs, const G16Fixed& fx, void* proof) {
    Fr s = Fr::load((const char*)rs + 32);
    auto jac1 = [](const char* p) {   // normalised Jacobian (x, y, 1) or (0, 1, 0)
        Fq z = Fq::load(p + 64);
        if (z.is_zero()) return G1x::inf();
        return G1x::from_affine(Fq::load(p), Fq::load(p + 32));
    };
    auto jac2 = [](const char* p) {
        host::Fq2 z = host::Fq2::load(p + 128);
        if (z.is_zero()) return G2x::inf();
        return G2x::from_affine(host::Fq2::load(p), host::Fq2::load(p + 64));
    };
    G1x sA = jac1(res), merged = jac1(res + 96);
    G2x sB2 = jac2(res + 192);
    uint64_t sc[4];
    s.canonical(sc);
    G1x A = g->alpha1.add(sA).add(fx.r_delta1);
    G2x B = g->beta2.add(sB2).add(fx.s_delta2);
    G1x C = merged.add(A.mul(sc)).add(fx.r_beta1);
    char* out = (char*)proof;
    A.store_affine(out);
    B.store_affine(out + 64);
    C.store_affine(out + 192);
}

}  // namespace hzk

extern "C" {

int hzk_g16_setup(hzk_ctx* ctx, uint64_t seed, int log_m, int num_public, hzk_g16** out) {
    if (!ctx || !out) return HZK_ERR_INVALID;
    *out = nullptr;
    HZK_CHECK(ctx, log_m >= 4 && log_m <= 24, HZK_ERR_INVALID, "g16_setup: log_m %d out of [4, 24]", log_m);
    HZK_CHECK(ctx, num_public >= 0 && num_public <= NUM_INPUTS, HZK_ERR_INVALID, "g16_setup: num_public %d", num_public);
    hzk_g16* g = new hzk_g16();
    g->log_m = log_m;
    g->M = (size_t)1 << log_m;
    g->num_public = num_public;
    g->num_gates = g->M - (num_public + 1);
    g->num_inputs = NUM_INPUTS;
    g->num_vars = 1 + NUM_INPUTS + g->num_gates;
    g->device = ctx->device;
    int rc = setup(ctx, g, seed);
    if (rc) {
        hzk_g16_free(g);
        return rc;
    }
    *out = g;
    return HZK_OK;
}

// The reference's MatMul statement (Groth16(input_scale) + Setup,
// groth16.cuh:39-71): k = 3 p^2 public inputs, N = 3 p^2 + p^3 variables
// besides z_0, M = next_pow2(p^3 + p^2) rows.
int hzk_g16_setup_matmul(hzk_ctx* ctx, uint64_t seed, int p, hzk_g16** out) {
    if (!ctx || !out) return HZK_ERR_INVALID;
    *out = nullptr;
    HZK_CHECK(ctx, p >= 2 && p <= 255, HZK_ERR_INVALID, "g16_setup_matmul: p %d out of [2, 255]", p);
    const size_t pp = (size_t)p, rows = pp * pp * pp + pp * pp;
    int log_m = 0;
    while (((size_t)1 << log_m) < rows) ++log_m;
    log_m = std::max(log_m, 4);
    HZK_CHECK(ctx, log_m <= 24, HZK_ERR_INVALID, "g16_setup_matmul: p %d needs a domain above 2^24", p);
    hzk_g16* g = new hzk_g16();
    g->kind = 1;
    g->mm_p = p;
    g->log_m = log_m;
    g->M = (size_t)1 << log_m;
    g->num_public = 3 * pp * pp;
    g->num_gates = rows;
    g->num_inputs = 2 * pp * pp;
    g->num_vars = 1 + 3 * pp * pp + pp * pp * pp;
    g->device = ctx->device;
    int rc = setup(ctx, g, seed);
    if (rc) {
        hzk_g16_free(g);
        return rc;
    }
    *out = g;
    return HZK_OK;
}

void hzk_g16_free(hzk_g16* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    for (void* p : g->owned) (void)hipFree(p);
    delete g;
}

int hzk_g16_info(const hzk_g16* g, uint64_t* out8) {
    if (!g || !out8) return HZK_ERR_INVALID;
    uint64_t kb = 0;   // every device array of the key: CSR, full and compacted query arrays, level tables, buffers
    for (size_t b : g->owned_bytes) kb += b;
    uint64_t v[8] = {g->M, g->num_vars, g->num_public, g->num_gates, g->m[0].col.size(), g->m[1].col.size(),
                     g->m[2].col.size(), kb};
    std::memcpy(out8, v, sizeof v);
    return HZK_OK;
}

int hzk_g16_witness(const hzk_g16* g, uint64_t wseed, void* z_host) {
    if (!g || !z_host) return HZK_ERR_INVALID;
    Fr* z = (Fr*)z_host;
    std::memset((void*)z, 0, g->num_vars * 32);
    z[0] = Fr::one();
    synth::Xoshiro rng(wseed, 3);
    for (size_t i = 1; i <= g->num_inputs; ++i) z[i] = fr_draw(rng);
    // forward: row r defines the variable of its single C entry (coefficient one)
    for (size_t r = 0; r < g->num_gates; ++r)
        z[g->m[2].col[g->m[2].row_ptr[r]]] = row_dot(g->m[0], g->M, r, z) * row_dot(g->m[1], g->M, r, z);
    return HZK_OK;
}

int hzk_g16_rs(uint64_t wseed, void* rs_host) {
    if (!rs_host) return HZK_ERR_INVALID;
    synth::Xoshiro rng(wseed, 5);
    Fr r = fr_draw(rng), s = fr_draw(rng);
    std::memcpy(rs_host, r.v, 32);
    std::memcpy((char*)rs_host + 32, s.v, 32);
    return HZK_OK;
}

int hzk_g16_prove(hzk_ctx* ctx, hzk_g16* g, const void* z_host, const void* rs_host, void* proof_host) {
    if (!ctx || !g) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, z_host && rs_host && proof_host, HZK_ERR_INVALID, "g16_prove: null argument");
    HZK_HIP(ctx, hipMemcpyAsync(g->z, z_host, g->num_vars * 32, hipMemcpyHostToDevice, ctx->stream));
    G16Req r{g, g->z, g->a, g->b, g->c, rs_host, {}};
    G16Fixed fx;
    std::thread fixed([&] { fx = g16_fixed_terms(g, rs_host); });   // host work beside the device's
    int rc = g16_h_batch(ctx, &r, 1);
    if (!rc) rc = g16_msm_batch(ctx, &r, 1, 0);
    fixed.join();
    if (rc) return rc;
    g16_assemble(g, r.res, rs_host, fx, proof_host);
    return HZK_OK;
}

// Prove over fixed-base level tables with `levels` levels per bucket set
// (built here, on ctx), or over the plain key arrays (levels <= 1).
int hzk_g16_set_levels(hzk_ctx* ctx, hzk_g16* g, int levels) {
    if (!ctx || !g) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, levels >= 0 && levels <= 64, HZK_ERR_INVALID, "g16_set_levels: %d", levels);
    return build_levels(ctx, g, (uint32_t)levels);
}

// Device pointer and element count of a proving-key array (tests, serving):
// which 0 A1, 1 B1, 2 B2 (G2), 3 L1, 4 H1.
int hzk_g16_key_array(const hzk_g16* g, int which, void** dev_ptr, size_t* count) {
    if (!g || !dev_ptr || !count) return HZK_ERR_INVALID;
    switch (which) {
        case 0: *dev_ptr = g->A1; *count = g->num_vars; break;
        case 1: *dev_ptr = g->B1; *count = g->num_vars; break;
        case 2: *dev_ptr = g->B2; *count = g->num_vars; break;
        case 3: *dev_ptr = g->L1; *count = g->num_vars - g->num_public - 1; break;
        case 4: *dev_ptr = g->H1; *count = g->M - 1; break;
        default: return HZK_ERR_INVALID;
    }
    return HZK_OK;
}

}  // extern "C"
