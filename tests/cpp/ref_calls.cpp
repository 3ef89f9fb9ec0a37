// The reference's own test programs, statement for statement, compiled with a
// host compiler against include/ and linked with libhzk.so:
//   test/msm-check.cu:17-41     G1: fixed-base table, pre_comp levels, MSM
//   test/msm-g2-check.cu:24-51  G2: the same with the generator record dev_uni
//   test/groth16.cu:5-8         Groth16_BN254 on input_scale::MatMul_p_100
// plus ntt_setup (include/protocol/groth16-operator.cuh:6-32) and the MatMul
// CSR of xpu::spmat (include/data-structure/xpu-spmat.cuh:33-60).
//
// Kernels the reference launches with <<<grid, block>>> are host functions
// of the same name here, so each such call keeps its parenthesised
// template-id and argument list and drops the launch configuration;
// CUDA_DEBUG lines and the Timer are left out. The reference then checks the
// MSM with libff on the CPU (msm-check.cu:43-57); here the inputs and outputs
// go to files that tests/test_gpu_ref_calls.py compares with the oracle.
//
//   ref_calls <dir> <log_n> <p>
//     writes k_g1.bin s_g1.bin lvl0_g1.bin dest_g1.bin (G1, n = 2^log_n),
//            k_g2.bin s_g2.bin lvl0_g2.bin dest_g2.bin (G2),
//            ntt_params.bin (M = 2^10), spmat_{a,b,c}_{col,rp}.bin (p = 3),
//            g16.txt + g16_proof.bin (MatMul p, 39 / 63 / 100), status.txt
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include "alt_bn128_types.hpp"
#include "protocol/groth16-operator.hpp"
#include "protocol/groth16.hpp"

namespace {
std::string dir;
size_t n = 1 << 20;
const size_t win_siz = 12;
const size_t win_cnt = (256 + win_siz - 1) / win_siz;
const size_t bucket_cnt = 1 << win_siz;

void spill(const char* name, const void* p, size_t bytes) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write((const char*)p, (std::streamsize)bytes);
}

void msm_g1_check() {
    libff::init_alt_bn128_params();

    xpu::vector<libff::alt_bn128_G1> dev_points(win_cnt * n, xpu::mem_policy::device_only);
    xpu::vector<libff::alt_bn128_Fr> dev_scalars(n, xpu::mem_policy::cross_platform);

    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) dev_scalars[i] = libff::alt_bn128_Fr::random_element();
    dev_scalars.store();
    spill("k_g1.bin", &dev_scalars[0], 32 * n);

    (fix_base_multi_scalar_multiplication_g1<alt_bn128::fr_t, alt_bn128::g1_t>)((alt_bn128::g1_t*)dev_points.p(), (alt_bn128::fr_t*)dev_scalars.p(), n);
    (pre_comp_g1<alt_bn128::g1_t>)(n, (alt_bn128::g1_t*)dev_points.p(), win_siz, win_cnt);

    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) dev_scalars[i] = libff::alt_bn128_Fr::random_element();
    dev_scalars.store();
    spill("s_g1.bin", &dev_scalars[0], 32 * n);

    multi_scalar_multiplication_g1<alt_bn128::fr_t, alt_bn128::g1_t>(n, (alt_bn128::fr_t*)dev_scalars.p(), (alt_bn128::g1_t*)dev_points.p(), (alt_bn128::g1_t*)dev_points.p() + n, win_siz, win_cnt, bucket_cnt, n);

    xpu::vector<libff::alt_bn128_G1> points(n+1);
    points.load(dev_points.p(), n+1);
    spill("lvl0_g1.bin", &points[0], 96 * n);
    spill("dest_g1.bin", &points[n], 96);
}

void msm_g2_check() {
    libff::init_alt_bn128_params();

    xpu::vector<libff::alt_bn128_G2> dev_points(win_cnt * n, xpu::mem_policy::device_only);
    xpu::vector<libff::alt_bn128_Fr> dev_scalars(n, xpu::mem_policy::cross_platform);

    xpu::vector<libff::alt_bn128_G2> dev_uni(1, xpu::mem_policy::cross_platform);
    dev_uni[0] = libff::alt_bn128_G2::one();
    dev_uni.store();

    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) dev_scalars[i] = libff::alt_bn128_Fr::random_element();
    dev_scalars.store();
    spill("k_g2.bin", &dev_scalars[0], 32 * n);

    (fix_base_multi_scalar_multiplication_g2<alt_bn128::fr_t, alt_bn128::g2_t>)((alt_bn128::g2_t*)dev_points.p(), (alt_bn128::fr_t*)dev_scalars.p(), n, (alt_bn128::g2_t*)dev_uni.p());
    (pre_comp_g2<alt_bn128::g2_t>)(n, (alt_bn128::g2_t*)dev_points.p(), win_siz, win_cnt);

    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) dev_scalars[i] = libff::alt_bn128_Fr::random_element();
    dev_scalars.store();
    spill("s_g2.bin", &dev_scalars[0], 32 * n);

    multi_scalar_multiplication_g2<alt_bn128::fr_t, alt_bn128::g2_t>(n, (alt_bn128::fr_t*)dev_scalars.p(), (alt_bn128::g2_t*)dev_points.p(), (alt_bn128::g2_t*)dev_points.p() + 2*n, win_siz, win_cnt, bucket_cnt, n);

    // the reference reorders the lane records with a kernel of its test
    // (layout_switch) before loading; the raw records are compared here
    xpu::vector<libff::alt_bn128_G2> points(n+1, xpu::mem_policy::cross_platform);
    points.load(dev_points.p(), n+1);
    spill("lvl0_g2.bin", &points[0], 192 * n);
    spill("dest_g2.bin", &points[n], 192);
}

void ntt_setup_check() {
    const size_t M = 1 << 10;
    xpu::vector<libff::alt_bn128_Fr> ntt_params;
    ntt_setup<libff::alt_bn128_Fr, &alt_bn128_ntt_config>(M, ntt_params);
    spill("ntt_params.bin", &ntt_params[0], 32 * (6 * M + 2));
}

void spmat_check() {
    xpu::spmat<libff::alt_bn128_Fr> A, B, C;
    A.init(xpu::application_scence::MatMul, xpu::matrix_type::A, 3);
    B.init(xpu::application_scence::MatMul, xpu::matrix_type::B, 3);
    C.init(xpu::application_scence::MatMul, xpu::matrix_type::C, 3);
    xpu::spmat<libff::alt_bn128_Fr>* m[3] = {&A, &B, &C};
    const char* names[3] = {"a", "b", "c"};
    for (int i = 0; i < 3; ++i) {
        auto d = m[i]->template descript<alt_bn128::fr_t>();
        std::string c = std::string("spmat_") + names[i] + "_col.bin", r = std::string("spmat_") + names[i] + "_rp.bin";
        spill(c.c_str(), &m[i]->host_col_idx()[0], 8 * d.cnt);
        spill(r.c_str(), &m[i]->host_row_ptr()[0], 8 * d.row);
    }
}

void groth16_check(int p) {
    libff::init_alt_bn128_params();
    input_scale is = p == 39 ? input_scale::MatMul_p_39 : (p == 63 ? input_scale::MatMul_p_63 : input_scale::MatMul_p_100);
    Groth16_BN254 protocol(is);
    protocol.Setup(12);
    protocol.Prove();
    spill("g16_proof.bin", protocol.proof().data(), 256);
    std::ofstream(dir + "/g16.txt") << p << " " << protocol.seed() << " " << protocol.last_witness_seed() << "\n";
}
}  // namespace

int main(int argc, char* argv[]) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: ref_calls <dir> <log_n> <p>\n");
        return 2;
    }
    dir = argv[1];
    n = (size_t)1 << std::atoi(argv[2]);
    int p = std::atoi(argv[3]);
    try {
        std::printf("ref_calls: G1 MSM n=%zu\n", n), std::fflush(stdout);
        msm_g1_check();
        std::printf("ref_calls: G2 MSM n=%zu\n", n), std::fflush(stdout);
        msm_g2_check();
        ntt_setup_check();
        spmat_check();
        std::printf("ref_calls: Groth16 MatMul p=%d\n", p), std::fflush(stdout);
        groth16_check(p);
        std::ofstream(dir + "/status.txt") << "ok\n";
    } catch (const lightzk::error& e) {
        std::ofstream(dir + "/status.txt") << "error " << e.code << "\n" << e.what() << "\n";
        std::fprintf(stderr, "ref_calls: %s\n", e.what());
        return 1;
    }
    return 0;
}
