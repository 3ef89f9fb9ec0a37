// Drop-in for the reference's protocol class (include/protocol/groth16.cuh:
// 11-229). input_scale keeps its names and its statement: the MatMul R1CS of
// xpu::spmat (include/data-structure/xpu-spmat.cuh:33-65) for p = 39, 63,
// 100 (M = 2^16, 2^18, 2^20; k = 3p^2 public inputs, N = 3p^2 + p^3).
// Setup(win_siz) builds the circuit and the proving key from seeded toxic
// waste, resident in HBM; Prove() proves a satisfying witness (seeded X and
// Y, their product computed forward; the reference draws a random,
// unsatisfying z, groth16.cuh:124-130) and keeps the proof (A | B | C
// affine, 256 bytes) instead of printing it (groth16.cuh:219-221). Setup's
// window size is a tuning hint: results do not depend on it.
#pragma once
#include <array>
#include <cstdint>
#include <vector>
#include "../lightzk_context.hpp"

enum class input_scale { MatMul_p_39, MatMul_p_63, MatMul_p_100 };

class Groth16 {
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

private:
    uint64_t seed_;
    int p_;
    hzk_g16* g_ = nullptr;
    size_t num_vars_ = 0;
    uint64_t wseed_ = 0;
    std::array<uint8_t, 256> proof_{};
};
