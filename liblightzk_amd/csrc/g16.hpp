// Groth16 prover internals shared by the plain prover (groth16.cpp) and the
// serving loop (server.cpp).
#pragma once
#include <vector>
#include "common.hpp"
#include "host_ec.hpp"
#include "keyfile_format.hpp"
#include "msm.hpp"

struct hzk_g16 {
    int log_m = 0;
    int kind = 0;   // 0 multiplication-gate DAG, 1 the reference's MatMul circuit (p = mm_p)
    int mm_p = 0;
    // rows 0..num_gates-1 each define the variable of their single C entry;
    // z[1..num_inputs] are the seeded inputs
    size_t M = 0, num_vars = 0, num_public = 0, num_gates = 0, num_inputs = 0;
    hzk::HostCsr m[3];
    // device copies
    void* d_val[3] = {nullptr, nullptr, nullptr};
    uint64_t* d_col[3] = {nullptr, nullptr, nullptr};
    uint64_t* d_rp[3] = {nullptr, nullptr, nullptr};
    // proving key in HBM (affine)
    void *A1 = nullptr, *B1 = nullptr, *B2 = nullptr, *L1 = nullptr, *H1 = nullptr;
    hzk::host::G1x alpha1, beta1, delta1;
    hzk::host::G2x beta2, gamma2, delta2;
    // byte-window tables of delta1, beta1, delta2 for the proofs' r delta1,
    // r beta1 and s delta2 (built with the prover buffers)
    hzk::host::FixedBase<hzk::host::Fq> fb_delta1, fb_beta1;
    hzk::host::FixedBase<hzk::host::Fq2> fb_delta2;
    std::vector<hzk::host::G1x> IC;   // verifying-key input commitments
    // buffers of the plain one-proof-at-a-time prover, and 1/(g^M - 1)
    void *z = nullptr, *a = nullptr, *b = nullptr, *c = nullptr, *zinv = nullptr;
    // The A and B query points of variables absent from A (resp. B) are the
    // point at infinity (a third of them for the gate circuit, most of B for
    // MatMul). The MSMs run over compacted copies holding only the finite
    // points: idxA / idxB list their variables (at least one entry each).
    // B1 enters the proof only through C = sum L + sum H + s A + r B1 - r s d1
    //   = [sum_{i>l} z_i L_i + sum_j h_j H_j + sum_{i in B} (r z_i) B1_i] + s A + r beta1,
    // so L, H and the compacted B1 form one MSM over C1 = L1 | H1 | B1c
    // (nC = nL + M - 1 + nB points). A proof's b buffer (free once h is
    // computed) takes z_A | z_B (vec_b elements), its c buffer
    // z_{l+1..} | h | r z_B (vec_c elements).
    uint32_t *idxA = nullptr, *idxB = nullptr;
    size_t nA = 0, nB = 0, nC = 0, vec_b = 0, vec_c = 0;
    void *A1c = nullptr, *B2c = nullptr, *C1 = nullptr;
    // A1c, C1 and B2c in the form the bucket accumulations read
    // (msm_convert_g1 / _g2), converted once here instead of in every MSM call
    void *A1cv = nullptr, *C1v = nullptr, *B2cv = nullptr;
    // fixed-base level tables of the key arrays for window size lc with lv
    // levels (lv <= 1: none); each MSM bucket set then serves lv windows
    // (held in the converted form)
    void *A1L = nullptr, *C1L = nullptr, *B2L = nullptr;
    int lc = 0;
    uint32_t lv = 1;
    std::vector<void*> owned;        // device allocations freed with the key
    std::vector<size_t> owned_bytes; // their sizes (resident key bytes = the sum)
    int device = 0;
    void own(void* p, size_t bytes) { owned.push_back(p), owned_bytes.push_back(bytes); }
    void disown(void* p) {
        for (size_t i = 0; i < owned.size(); ++i)
            if (owned[i] == p) {
                owned.erase(owned.begin() + (long)i), owned_bytes.erase(owned_bytes.begin() + (long)i);
                return;
            }
    }
};

namespace hzk {

// One proof in flight: device buffers for z and the three evaluation vectors
// (a holds h after the H stage), the five MSM results, randomness and output.
struct G16Req {
    hzk_g16* g;
    void *z, *a, *b, *c;
    const void* rs;  // r | s (host, Montgomery)
    char res[384];   // sum A (G1, 96 B), the merged C MSM (G1, 96 B), sum B2 (G2, 192 B)
};
// SpMV and the H pipeline (3 iNTT, 3 coset NTT, (AB - C)/Z, coset iNTT) for
// several requests, enqueued on ctx's stream with batched transforms.
int g16_h_batch(hzk_ctx* ctx, G16Req* reqs, size_t n);
// the MSMs of several requests as two batched MSMs (G1: A and the merged C
// MSM per request, G2: B per request); results land in reqs[i].res on the host
int g16_msm_batch(hzk_ctx* ctx, G16Req* reqs, size_t n, int graph);
// the proof terms that depend only on the key and r, s: r delta1, s delta2,
// r beta1 (host scalar multiplications, computed while the device proves)
struct G16Fixed {
    host::G1x r_delta1, r_beta1;
    host::G2x s_delta2;
};
G16Fixed g16_fixed_terms(const hzk_g16* g, const void* rs);
// phase-6 assembly on the host; proof = A | B | C (affine, 256 B)
void g16_assemble(const hzk_g16* g, const char* res, const void* rs, const G16Fixed& fx, void* proof);
// device bytes one request needs beyond its buffers (MSM scratch, generous)
size_t g16_request_footprint(const hzk_g16* g);
// allocate the plain prover's buffers (z, a, b, c) and upload 1/(g^M - 1)
int g16_prover_buffers(hzk_ctx* ctx, hzk_g16* g);
// compacted A / B query arrays and their index lists from the full key arrays
int g16_sparse_key(hzk_ctx* ctx, hzk_g16* g);

}  // namespace hzk
