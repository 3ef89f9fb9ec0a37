// Exercises the reference-shaped C++ API (include/operator/*.hpp,
// include/protocol/groth16.hpp) with explicit template arguments, the way the
// reference's tests call the operators (test/msm-check.cu:28-39,
// test/msm-g2-check.cu:38-49, test/groth16.cu:6-8), on
// inputs written by tests/test_gpu_shims.py, and writes the outputs back for
// that test to compare with the oracle and the C ABI.
//
//   shim_check <dir>   reads  params.txt scalars.bin ks.bin ntt_in.bin ntt_tab.bin
//                             ntt_factor.bin csr_{a,b,c}_{val,col,rp}.bin z.bin
//                      writes msm_g1.bin msm_g2.bin levels_g1.bin levels_g2.bin
//                             ntt_out.bin spmv.bin abc.bin g16_proof.bin status.txt
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>
#include "alt_bn128_types.hpp"
#include "operator/msm.hpp"
#include "operator/poly.hpp"
#include "protocol/groth16.hpp"

namespace {

std::string dir;

std::vector<uint8_t> slurp(const char* name) {
    std::ifstream f(dir + "/" + name, std::ios::binary);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
void spill(const char* name, const std::vector<uint8_t>& v) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write((const char*)v.data(), (std::streamsize)v.size());
}

// device buffer through the C ABI (the reference uses xpu::vector for this)
struct dev {
    void* p = nullptr;
    size_t n = 0;
    explicit dev(size_t bytes) : n(bytes) { lightzk::check(hzk_device_alloc(lightzk::context(), bytes ? bytes : 1, &p)); }
    explicit dev(const std::vector<uint8_t>& h) : dev(h.size()) { put(h); }
    ~dev() { hzk_device_free(lightzk::context(), p); }
    void put(const std::vector<uint8_t>& h) { lightzk::check(hzk_memcpy_h2d(lightzk::context(), p, h.data(), h.size())); }
    std::vector<uint8_t> get(size_t bytes, size_t off = 0) const {
        std::vector<uint8_t> h(bytes);
        lightzk::check(hzk_memcpy_d2h(lightzk::context(), h.data(), (const char*)p + off, bytes));
        return h;
    }
};

typedef alt_bn128::fr_t fr;
typedef alt_bn128::g1_t g1rec;   // the reference's 96-byte device G1 record
typedef alt_bn128::g2_t g2rec;   // one lane record of a G2 point

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) { std::fprintf(stderr, "usage: shim_check <dir>\n"); return 2; }
    dir = argv[1];
    size_t n = 0, gap = 0, level = 0, deg = 0, M = 0, nv = 0, nnz[3] = {0, 0, 0};
    {
        std::ifstream p(dir + "/params.txt");
        p >> n >> gap >> level >> deg >> M >> nv >> nnz[0] >> nnz[1] >> nnz[2];
    }
    try {
        // MSM (test/msm-check.cu:28-39 order: fixed-base table, pre_comp levels, MSM)
        dev ks(slurp("ks.bin")), sc(slurp("scalars.bin"));
        dev t1(96 * n * level), d1(96);
        (fix_base_multi_scalar_multiplication_g1<fr, g1rec>)((g1rec*)t1.p, (fr*)ks.p, n);
        (pre_comp_g1<g1rec>)(n, (g1rec*)t1.p, gap, level);
        multi_scalar_multiplication_g1<fr, g1rec>(n, (fr*)sc.p, (g1rec*)t1.p, (g1rec*)d1.p, 12, 22, 4096, n);
        spill("msm_g1.bin", d1.get(96));
        {
            std::vector<uint8_t> lv;
            for (size_t j = 0; j < level; ++j) {
                auto r = t1.get(96, 96 * j * n);
                lv.insert(lv.end(), r.begin(), r.end());
            }
            spill("levels_g1.bin", lv);
        }
        dev t2(192 * n * level), d2(192);
        (fix_base_multi_scalar_multiplication_g2<fr, g2rec>)((g2rec*)t2.p, (fr*)ks.p, n);
        (pre_comp_g2<g2rec>)(n, (g2rec*)t2.p, gap, level);
        multi_scalar_multiplication_g2<fr, g2rec>(n, (fr*)sc.p, (g2rec*)t2.p, (g2rec*)d2.p, 12, 22, 4096, n);
        spill("msm_g2.bin", d2.get(192));
        {
            std::vector<uint8_t> lv;
            for (size_t j = 0; j < level; ++j) {
                auto r = t2.get(192, 192 * j * n);
                lv.insert(lv.end(), r.begin(), r.end());
            }
            spill("levels_g2.bin", lv);
        }
        // NTT with the caller's table and factor (groth16.cuh:166-168 call shape)
        dev x(slurp("ntt_in.bin")), tab(slurp("ntt_tab.bin")), fac(slurp("ntt_factor.bin"));
        int log_deg = 0;
        while (((size_t)1 << log_deg) < deg) ++log_deg;
        number_theory_transformation<fr>(deg, (size_t)log_deg, (fr*)x.p, (fr*)tab.p, (fr*)fac.p);
        spill("ntt_out.bin", x.get(32 * deg));
        // SpMV and A o B - C (groth16.cuh:160-162, :188)
        const char* mats[3] = {"a", "b", "c"};
        std::vector<dev*> keep;
        xpu::spmat_descript<fr> d[3];
        for (int m = 0; m < 3; ++m) {
            std::string pre = std::string("csr_") + mats[m] + "_";
            dev* v = new dev(slurp((pre + "val.bin").c_str()));
            dev* c = new dev(slurp((pre + "col.bin").c_str()));
            dev* r = new dev(slurp((pre + "rp.bin").c_str()));
            keep.insert(keep.end(), {v, c, r});
            d[m] = {(fr*)v->p, (size_t*)c->p, (size_t*)r->p, M, nv, nnz[m]};
        }
        dev z(slurp("z.bin")), az(32 * M), bz(32 * M), cz(32 * M);
        (batch_sparse_matrix_vector_multiplication<fr>)(d[0], d[1], d[2], (fr*)z.p, (fr*)az.p, (fr*)bz.p, (fr*)cz.p);
        std::vector<uint8_t> out = az.get(32 * M), tb = bz.get(32 * M), tc = cz.get(32 * M);
        out.insert(out.end(), tb.begin(), tb.end());
        out.insert(out.end(), tc.begin(), tc.end());
        spill("spmv.bin", out);
        (polynomial_A_times_B_minus_C<fr>)(M, (fr*)az.p, (fr*)bz.p, (fr*)cz.p);
        spill("abc.bin", az.get(32 * M));
        for (dev* p : keep) delete p;
        // protocol class (test/groth16.cu:3-8)
        Groth16_BN254 g(input_scale::MatMul_p_39);
        g.Setup(12);
        g.Prove(1);
        spill("g16_proof.bin", std::vector<uint8_t>(g.proof().begin(), g.proof().end()));
        // error behaviour: an exception carrying the library's message
        std::string err = "none";
        try {
            number_theory_transformation<fr>(12, 3, (fr*)x.p, (fr*)tab.p, (fr*)nullptr);
        } catch (const lightzk::error& e) {
            err = e.what();
        }
        std::ofstream(dir + "/status.txt") << "ok\n" << err << "\n";
    } catch (const lightzk::error& e) {
        std::ofstream(dir + "/status.txt") << "error " << e.code << "\n" << e.what() << "\n";
        std::fprintf(stderr, "shim_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
