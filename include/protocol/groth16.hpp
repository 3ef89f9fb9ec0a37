// Drop-in for the reference's protocol class (include/protocol/groth16.cuh:
// 11-229). input_scale keeps its names and domain sizes (M = 2^16, 2^18,
// 2^20); the statement is this library's seeded multiplication-gate circuit
// of that domain size (the reference's MatMul generator is not part of this
// library). Setup(win_siz) builds the circuit and the proving key, resident
// in HBM; Prove() proves a witness drawn from the next seed and keeps the
// proof (A | B | C affine, 256 bytes) instead of printing it
// (groth16.cuh:219-221). Setup's window size is a tuning hint: results do not
// depend on it.
#pragma once
#include <array>
#include <cstdint>
#include <vector>
#include "../lightzk_context.hpp"

enum class input_scale { MatMul_p_39, MatMul_p_63, MatMul_p_100 };

class Groth16 {
public:
    explicit Groth16(input_scale is, uint64_t seed = 2) : seed_(seed) {
        log_m_ = is == input_scale::MatMul_p_39 ? 16 : (is == input_scale::MatMul_p_63 ? 18 : 20);
    }
    ~Groth16() { hzk_g16_free(g_); }
    Groth16(const Groth16&) = delete;
    Groth16& operator=(const Groth16&) = delete;

    void Setup(size_t win_siz) {
        (void)win_siz;
        hzk_g16_free(g_);
        g_ = nullptr;
        lightzk::check(hzk_g16_setup(lightzk::context(), seed_, log_m_, 8, &g_));
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
    int log_m_;
    hzk_g16* g_ = nullptr;
    size_t num_vars_ = 0;
    uint64_t wseed_ = 0;
    std::array<uint8_t, 256> proof_{};
};
