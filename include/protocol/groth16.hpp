// Drop-in for the reference's protocol class (include/protocol/groth16.cuh:
// 11-229): the class template Groth16<FieldT, G1T, G2T, devFdT, dG1, dG2,
// ntt_config*> with input_scale, Setup(win_siz), Prove(), and the
// Groth16_BN254 instantiation. input_scale keeps its names and its
// statement: the MatMul R1CS of xpu::spmat
// (include/data-structure/xpu-spmat.cuh:33-65) for p = 39, 63, 100 (M = 2^16,
// 2^18, 2^20; k = 3p^2 public inputs, N = 3p^2 + p^3).
//
// Setup(win_siz) builds the circuit and the proving key from seeded toxic
// waste, resident in HBM (the reference draws random key points,
// groth16.cuh:84-122). Prove() proves a satisfying witness (seeded X and Y,
// their product computed forward; the reference draws a random, unsatisfying
// z, groth16.cuh:124-130) and keeps the proof (A | B | C affine, 256 bytes)
// for proof() instead of printing it (groth16.cuh:219-221). Setup's window
// size is a tuning hint: results do not depend on it. The template arguments
// name the caller's types; their byte sizes are checked, nothing else of them
// is used.
#pragma once
#include <array>
#include <cstdint>
#include <vector>
#include "../alt_bn128_types.hpp"
#include "../lightzk_context.hpp"
#include "groth16-operator.hpp"

enum class input_scale { MatMul_p_39, MatMul_p_63, MatMul_p_100 };

template <typename FieldT, typename G1T, typename G2T, typename devFdT, typename dG1, typename dG2, ntt_config* nc>
class Groth16 {
    static_assert(sizeof(FieldT) == 32 && sizeof(devFdT) == 32, "Fr elements are 32 bytes");
    static_assert(sizeof(G1T) == 96 && sizeof(dG1) == 96 && sizeof(G2T) == 192 && sizeof(dG2) == 96,
                  "reference point layouts");

public:
    explicit Groth16(input_scale is, uint64_t seed = 2) : seed_(seed) {
        p_ = is == input_scale::MatMul_p_39 ? 39 : (is == input_scale::MatMul_p_63 ? 63 : 100);
    }
    ~Groth16() { hzk_g16_free(g_); }
    Groth16(const Groth16&) = delete;
    Groth16& operator=(const Groth16&) = delete;

    void Setup(size_t win_siz) {
        (void)win_siz;
        hzk_g16_free(g_);
        g_ = nullptr;
        lightzk::check(hzk_g16_setup_matmul(lightzk::context(), seed_, p_, &g_));
        uint64_t info[8];
        hzk_g16_info(g_, info);
        num_vars_ = info[1];
    }

    void Prove() { Prove(++wseed_); }
    void Prove(uint64_t witness_seed) {
        if (!g_) throw lightzk::error(HZK_ERR_INVALID, "Groth16::Prove before Setup");
        std::vector<uint8_t> z(num_vars_ * 32);
        uint8_t rs[64];
        lightzk::check(hzk_g16_witness(g_, witness_seed, z.data()));
        lightzk::check(hzk_g16_rs(witness_seed, rs));
        lightzk::check(hzk_g16_prove(lightzk::context(), g_, z.data(), rs, proof_.data()));
    }

    const std::array<uint8_t, 256>& proof() const { return proof_; }
    hzk_g16* handle() const { return g_; }
    size_t num_vars() const { return num_vars_; }
    uint64_t seed() const { return seed_; }
    uint64_t last_witness_seed() const { return wseed_; }

private:
    uint64_t seed_;
    int p_;
    hzk_g16* g_ = nullptr;
    size_t num_vars_ = 0;
    uint64_t wseed_ = 0;
    std::array<uint8_t, 256> proof_{};
};

// include/protocol/groth16.cuh:229
typedef Groth16<libff::alt_bn128_Fr, libff::alt_bn128_G1, libff::alt_bn128_G2, alt_bn128::fr_t, alt_bn128::g1_t,
                alt_bn128::g2_t, &alt_bn128_ntt_config>
    Groth16_BN254;
