// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only,
// no GPU, no HIP): the product's host arithmetic (liblightzk_amd/csrc/
// host_ec.hpp), its seeded input streams (synth.cpp), the key-file header,
// R1CS and trailer checks (keyfile_format.hpp) on a valid file and on every
// corruption the loader must reject, and the oracle's CPU restatement
// (oracle/*.cpp: field, curve, MSM, NTT, pairing, Groth16), the host
// fixed-base tables (host_ec.hpp FixedBase) and the host worker pool
// (host_pool.cpp) under concurrent and nested callers. Values are compared
// with the oracle where both compute the same thing.
//
//   host_sanitize <scratch dir>     prints "host_sanitize ok" and exits 0
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include "../../liblightzk_amd/csrc/host_ec.hpp"
#include "../../liblightzk_amd/csrc/host_pool.hpp"
// the NTT's 29-bit Fr arithmetic compiled for the host: it needs only the
// eight-limb value type and a function qualifier
namespace hzk {
struct fr_t {
    uint32_t l[8];
};
}  // namespace hzk
#define HZK_DEV inline
#include "../../liblightzk_amd/csrc/fr29.hpp"
#include "../../liblightzk_amd/csrc/keyfile_format.hpp"

extern "C" {
int hzk_gen_field(uint64_t seed, uint64_t stream, size_t n, int kind, void* host_out);
int orc_fq_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n);
int orc_fr_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n);
int orc_fq2_op(int op, const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n);
int orc_g1_mul_gen(const uint64_t* k, uint64_t* out, size_t n);
int orc_g2_mul_gen(const uint64_t* k, uint64_t* out, size_t n);
int orc_gen_fr(uint64_t seed, uint64_t stream, size_t n, int sparse, uint64_t* out);
int orc_gen_fq(uint64_t seed, uint64_t stream, size_t n, uint64_t* out);
int orc_msm_g1(const uint64_t* s, const uint64_t* bases, size_t n, int method, uint64_t* out);
int orc_msm_g2(const uint64_t* s, const uint64_t* bases, size_t n, int method, uint64_t* out);
int orc_msm_closed_form(const uint64_t* s, const uint64_t* k, size_t n, int group, uint64_t* out);
int orc_ntt(size_t deg, size_t log_deg, uint64_t* a, const uint64_t* omega, const uint64_t* factor);
int orc_ntt_setup(size_t M, uint64_t* out);
int orc_pairing_product_is_one(const uint64_t* ps, const uint64_t* qs, size_t n);
int orc_circuit_sizes(uint64_t seed, int log_m, int num_public, uint64_t* out7);
int orc_circuit_csr(uint64_t seed, int log_m, int num_public, int which, uint64_t* row_ptr, uint64_t* col, uint64_t* val);
int orc_circuit_witness(uint64_t seed, int log_m, int num_public, uint64_t wseed, uint64_t* out);
void* orc_g16_new(uint64_t seed, int log_m, int num_public);
void orc_g16_free(void* h);
size_t orc_g16_key_points(void* h, int which, uint64_t* out);
void orc_g16_rs(uint64_t wseed, uint64_t* out);
int orc_g16_prove_trapdoor(void* h, const uint64_t* z, const uint64_t* rs, uint64_t* proof);
int orc_g16_prove(void* h, const uint64_t* z, const uint64_t* rs, uint64_t* proof);
int orc_g16_verify(void* h, const uint64_t* public_z, const uint64_t* proof);
}

using namespace hzk;
using U = std::vector<uint64_t>;

static int failures = 0;
#define CHECK(cond, what)                                                          \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, what);     \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

static void check_fields() {
    const size_t n = 512;
    U a(4 * n), b(4 * n), got(4 * n), want(4 * n);
    orc_gen_fq(71, 0, n, a.data());
    orc_gen_fq(71, 1, n, b.data());
    for (int op = 0; op < 3; ++op) {
        for (size_t i = 0; i < n; ++i) {
            host::Fq x = host::Fq::load(&a[4 * i]), y = host::Fq::load(&b[4 * i]);
            host::Fq r = op == 0 ? x * y : op == 1 ? x + y : x - y;
            r.store(&got[4 * i]);
        }
        orc_fq_op(op, a.data(), b.data(), want.data(), n);
        CHECK(got == want, "host Fq op vs oracle");
    }
    for (size_t i = 0; i < n; ++i) host::Fq::load(&a[4 * i]).inverse().store(&got[4 * i]);
    orc_fq_op(5, a.data(), nullptr, want.data(), n);
    CHECK(got == want, "host Fq inverse vs oracle");
    orc_gen_fr(71, 2, n, 0, a.data());
    orc_gen_fr(71, 3, n, 0, b.data());
    for (size_t i = 0; i < n; ++i) (host::Fr::load(&a[4 * i]) * host::Fr::load(&b[4 * i])).store(&got[4 * i]);
    orc_fr_op(0, a.data(), b.data(), want.data(), n);
    CHECK(got == want, "host Fr mul vs oracle");
    U a2(8 * n), b2(8 * n), g2(8 * n), w2(8 * n);
    orc_gen_fq(72, 0, 2 * n, a2.data());
    orc_gen_fq(72, 1, 2 * n, b2.data());
    for (size_t i = 0; i < n; ++i) (host::Fq2::load(&a2[8 * i]) * host::Fq2::load(&b2[8 * i])).store(&g2[8 * i]);
    orc_fq2_op(0, a2.data(), b2.data(), w2.data(), n);
    CHECK(g2 == w2, "host Fq2 mul vs oracle");
}

static void check_points_and_synth() {
    const size_t n = 64;
    U k(4 * n), aff1(8 * n), aff2(16 * n), got1(8 * n), got2(16 * n);
    CHECK(hzk_gen_field(73, 1, n, 0, k.data()) == 0, "hzk_gen_field");
    U ko(4 * n);
    orc_gen_fr(73, 1, n, 0, ko.data());
    CHECK(k == ko, "product and oracle streams agree");
    U sp(4 * 4096), spo(4 * 4096);
    hzk_gen_field(73, 2, 4096, 1, sp.data());
    orc_gen_fr(73, 2, 4096, 1, spo.data());
    CHECK(sp == spo, "sparse streams agree");
    orc_g1_mul_gen(k.data(), aff1.data(), n);
    orc_g2_mul_gen(k.data(), aff2.data(), n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t c[4];
        host::Fr::load(&k[4 * i]).canonical(c);
        host::g1_generator().mul(c).store_affine(&got1[8 * i]);
        host::g2_generator().mul(c).store_affine(&got2[16 * i]);
    }
    CHECK(got1 == aff1, "host G1 k*G vs oracle");
    CHECK(got2 == aff2, "host G2 k*G vs oracle");

    // byte-window tables of G: k*G by 32 table additions, k = 0 and k = 1 included
    host::FixedBase<host::Fq> fb1;
    host::FixedBase<host::Fq2> fb2;
    fb1.build(host::g1_generator());
    fb2.build(host::g2_generator());
    k[0] = k[1] = k[2] = k[3] = 0;
    k[4] = host::FR.one[0]; k[5] = host::FR.one[1]; k[6] = host::FR.one[2]; k[7] = host::FR.one[3];
    orc_g1_mul_gen(k.data(), aff1.data(), n);
    orc_g2_mul_gen(k.data(), aff2.data(), n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t c[4];
        host::Fr::load(&k[4 * i]).canonical(c);
        fb1.mul(c).store_affine(&got1[8 * i]);
        fb2.mul(c).store_affine(&got2[16 * i]);
    }
    CHECK(got1 == aff1, "host G1 fixed-base table vs oracle");
    CHECK(got2 == aff2, "host G2 fixed-base table vs oracle");
}

// fr29 (the NTT's two-element passes) against the oracle's Fr: the product
// of a value below 2^256 (brought below 2r) with a canonical twiddle scaled
// by 2^5 is the 2^256-form Montgomery product; sums and differences are
// congruent and stay below 2r; canonical() gives the oracle's bytes.
static void check_fr29() {
    const size_t n = 512;
    U a(4 * n), b(4 * n), prod(4 * n), sum(4 * n), dif(4 * n);
    orc_gen_fr(91, 1, n, 0, a.data());
    orc_gen_fr(91, 2, n, 0, b.data());
    // edge values: 0, 1 (Montgomery), r - 1 and, for a, the largest 2^256 - 1
    for (int q = 0; q < 4; ++q) {
        a[q] = b[q] = 0;
        a[4 + q] = b[4 + q] = host::FR.one[q];
        a[8 + q] = host::FR.p[q] - (q == 0);
        b[8 + q] = host::FR.p[q] - (q == 0);
        a[12 + q] = ~0ull;
    }
    U bc(b), ac(a);
    for (int q = 0; q < 4; ++q) ac[12 + q] = 0;   // the oracle's operands stay canonical
    orc_fr_op(0, ac.data(), bc.data(), prod.data(), n);
    orc_fr_op(1, ac.data(), bc.data(), sum.data(), n);
    orc_fr_op(2, ac.data(), bc.data(), dif.data(), n);
    for (size_t i = 0; i < n; ++i) {
        if (i == 3) continue;   // 2^256 - 1: checked below against the reduced value
        fr_t x, y;
        for (int w = 0; w < 4; ++w) {
            x.l[2 * w] = (uint32_t)a[4 * i + w], x.l[2 * w + 1] = (uint32_t)(a[4 * i + w] >> 32);
            y.l[2 * w] = (uint32_t)b[4 * i + w], y.l[2 * w + 1] = (uint32_t)(b[4 * i + w] >> 32);
        }
        const fr29 xa = fr29::unpack(x).reduce(), yb = fr29::unpack(y).reduce();
        const fr_t p = fr29::mul(xa, fr29::unpack(y).times32()).canonical().pack();
        const fr_t s = fr29::add(xa, yb).reduce().canonical().pack();
        const fr_t d = fr29::sub2(xa, yb).reduce().canonical().pack();
        for (int w = 0; w < 4; ++w) {
            CHECK((p.l[2 * w] | (uint64_t)p.l[2 * w + 1] << 32) == prod[4 * i + w], "fr29 product vs oracle");
            CHECK((s.l[2 * w] | (uint64_t)s.l[2 * w + 1] << 32) == sum[4 * i + w], "fr29 sum vs oracle");
            CHECK((d.l[2 * w] | (uint64_t)d.l[2 * w + 1] << 32) == dif[4 * i + w], "fr29 difference vs oracle");
        }
    }
    // 2^256 - 1 (not canonical) times one: reduce() then the product gives (2^256 - 1) mod r
    fr_t x, one;
    for (int w = 0; w < 8; ++w) x.l[w] = ~0u;
    for (int w = 0; w < 4; ++w) one.l[2 * w] = (uint32_t)host::FR.one[w], one.l[2 * w + 1] = (uint32_t)(host::FR.one[w] >> 32);
    const fr_t got = fr29::mul(fr29::unpack(x).reduce(), fr29::unpack(one).times32()).canonical().pack();
    // expected: (2^256 - 1) mod r, by repeated subtraction of r on 64-bit words
    uint64_t e[4] = {~0ull, ~0ull, ~0ull, ~0ull};
    for (;;) {
        bool ge = true;
        for (int w = 3; w >= 0; --w) {
            if (e[w] != host::FR.p[w]) { ge = e[w] > host::FR.p[w]; break; }
        }
        if (!ge) break;
        unsigned __int128 bw = 0;
        for (int w = 0; w < 4; ++w) {
            const unsigned __int128 v = (unsigned __int128)e[w] - host::FR.p[w] - (uint64_t)bw;
            e[w] = (uint64_t)v;
            bw = (v >> 64) ? 1 : 0;
        }
    }
    for (int w = 0; w < 4; ++w)
        CHECK((got.l[2 * w] | (uint64_t)got.l[2 * w + 1] << 32) == e[w], "fr29 reduce of 2^256 - 1");
}

// Four threads call parallel_for at once (as a server's pipelines do), with
// ragged lengths and a nested call inside some indices; every index runs once.
static void check_host_pool() {
    std::atomic<long> bad{0};
    std::vector<std::thread> callers;
    for (int t = 0; t < 4; ++t)
        callers.emplace_back([&bad, t] {
            for (int rep = 0; rep < 40; ++rep) {
                const long n = (rep * 37 + t * 11) % 97;   // 0 and 1 included
                std::vector<int> hit(n, 0);
                hzk::parallel_for(n, [&](long i) {
                    ++hit[i];
                    if (i % 16 == 0) {
                        std::atomic<int> inner{0};
                        hzk::parallel_for(5, [&](long) { ++inner; });
                        if (inner != 5) ++bad;
                    }
                });
                for (long i = 0; i < n; ++i)
                    if (hit[i] != 1) ++bad;
            }
        });
    for (auto& c : callers) c.join();
    CHECK(bad == 0, "host pool runs every index exactly once");
}

static void check_oracle_msm_ntt_pairing() {
    const size_t n = 300;
    U s(4 * n), k(4 * n), b1(8 * n), b2(16 * n), r1(12), r1n(12), c1(12), r2(24), c2(24);
    orc_gen_fr(74, 0, n, 0, s.data());
    orc_gen_fr(74, 1, n, 0, k.data());
    orc_g1_mul_gen(k.data(), b1.data(), n);
    orc_g2_mul_gen(k.data(), b2.data(), n);
    orc_msm_g1(s.data(), b1.data(), n, 1, r1.data());
    orc_msm_g1(s.data(), b1.data(), n, 0, r1n.data());
    orc_msm_closed_form(s.data(), k.data(), n, 1, c1.data());
    CHECK(r1 == c1 && r1n == c1, "oracle G1 MSM: Pippenger, naive, closed form");
    orc_msm_g2(s.data(), b2.data(), n, 1, r2.data());
    orc_msm_closed_form(s.data(), k.data(), n, 2, c2.data());
    CHECK(r2 == c2, "oracle G2 MSM vs closed form");
    const size_t M = 256;
    U tab(4 * (6 * M + 2)), x(4 * M), y;
    orc_ntt_setup(M, tab.data());
    orc_gen_fr(74, 2, M, 0, x.data());
    y = x;
    orc_ntt(M, 8, y.data(), tab.data(), nullptr);
    orc_ntt(M, 8, y.data(), tab.data() + 4 * M, tab.data() + 8 * M);
    CHECK(y == x, "oracle NTT round trip");
    // e(aP, bQ) e(-abP, Q) = 1
    U sc(8), pts(16), qs(32);
    orc_gen_fr(74, 3, 2, 0, sc.data());
    U ab(4);
    orc_fr_op(0, sc.data(), sc.data() + 4, ab.data(), 1);
    orc_g1_mul_gen(sc.data(), pts.data(), 1);
    U negab(4);
    orc_fr_op(3, ab.data(), nullptr, negab.data(), 1);
    orc_g1_mul_gen(negab.data(), pts.data() + 8, 1);
    orc_g2_mul_gen(sc.data() + 4, qs.data(), 1);
    U oneM(4);
    host::Fr::one().store(oneM.data());
    orc_g2_mul_gen(oneM.data(), qs.data() + 16, 1);
    CHECK(orc_pairing_product_is_one(pts.data(), qs.data(), 2) == 1, "pairing bilinearity");
}

static void check_groth16(int log_m) {
    void* h = orc_g16_new(9, log_m, 8);
    CHECK(h != nullptr, "oracle circuit");
    if (!h) return;
    uint64_t sz[7];
    orc_circuit_sizes(9, log_m, 8, sz);
    U z(4 * sz[1]), rs(8), p1(32), p2(32);
    orc_circuit_witness(9, log_m, 8, 3, z.data());
    orc_g16_rs(3, rs.data());
    orc_g16_prove(h, z.data(), rs.data(), p1.data());
    orc_g16_prove_trapdoor(h, z.data(), rs.data(), p2.data());
    CHECK(p1 == p2, "honest prover == trapdoor closed form");
    CHECK(orc_g16_verify(h, z.data(), p1.data()) == 1, "proof verifies");
    p1[0] ^= 1;
    CHECK(orc_g16_verify(h, z.data(), p1.data()) == 0, "tampered proof rejected");
    orc_g16_free(h);
}

// ---- key files: a valid file from the oracle's circuit, then corruptions
struct KeyImage {
    keyfile::Header h{};
    HostCsr m[3];
    U arrays;     // A1 B1 B2 L1 H1 as affine words
    U points;     // alpha1 beta1 delta1 beta2 gamma2 delta2 IC
};

static KeyImage make_image(int log_m) {
    KeyImage k;
    uint64_t sz[7];
    orc_circuit_sizes(9, log_m, 8, sz);
    keyfile::Header& h = k.h;
    std::memcpy(h.magic, keyfile::MAGIC, 8);
    h.version = keyfile::VERSION;
    h.kind = 0, h.mm_p = 0, h.log_m = (uint32_t)log_m;
    h.M = sz[0], h.num_vars = sz[1], h.num_public = sz[2], h.num_gates = sz[3], h.num_inputs = 64;
    for (int i = 0; i < 3; ++i) {
        h.nnz[i] = sz[4 + i];
        k.m[i].row_ptr.resize(h.M);
        k.m[i].col.resize(h.nnz[i]);
        k.m[i].val.resize(h.nnz[i]);
        orc_circuit_csr(9, log_m, 8, i, k.m[i].row_ptr.data(), k.m[i].col.data(), (uint64_t*)k.m[i].val.data());
    }
    void* o = orc_g16_new(9, log_m, 8);
    auto dump = [&](int which, size_t words) {
        size_t cnt = orc_g16_key_points(o, which, nullptr);
        U t(words * cnt);
        orc_g16_key_points(o, which, t.data());
        return t;
    };
    for (int which : {0, 1}) { U t = dump(which, 8); k.arrays.insert(k.arrays.end(), t.begin(), t.end()); }
    { U t = dump(2, 16); k.arrays.insert(k.arrays.end(), t.begin(), t.end()); }
    for (int which : {3, 4}) { U t = dump(which, 8); k.arrays.insert(k.arrays.end(), t.begin(), t.end()); }
    { U t = dump(6, 8); k.points.insert(k.points.end(), t.begin(), t.end()); }
    { U t = dump(7, 16); k.points.insert(k.points.end(), t.begin(), t.end()); }
    { U t = dump(5, 8); k.points.insert(k.points.end(), t.begin(), t.end()); }
    orc_g16_free(o);
    return k;
}

static void write_image(const std::string& path, const KeyImage& k, bool fix_checksum = true, long truncate = -1,
                        bool extra = false) {
    FILE* f = std::fopen(path.c_str(), "wb");
    keyfile::Writer w{f};
    w.put(&k.h, sizeof k.h);
    for (int i = 0; i < 3; ++i) {
        w.put(k.m[i].row_ptr.data(), k.m[i].row_ptr.size() * 8);
        w.put(k.m[i].col.data(), k.m[i].col.size() * 8);
        w.put(k.m[i].val.data(), k.m[i].val.size() * 32);
    }
    w.put(k.arrays.data(), k.arrays.size() * 8);
    w.put(k.points.data(), k.points.size() * 8);
    uint64_t cs = w.sum.done() ^ (fix_checksum ? 0 : 1);
    std::fwrite(&cs, 1, 8, f);
    if (extra) std::fputc(0, f);
    std::fclose(f);
    if (truncate >= 0) {
        FILE* g = std::fopen(path.c_str(), "rb");
        std::vector<char> all(1 << 20);
        size_t got = std::fread(all.data(), 1, all.size(), g);
        std::fclose(g);
        all.resize(got);
        g = std::fopen(path.c_str(), "wb");
        std::fwrite(all.data(), 1, std::min<size_t>(got, (size_t)truncate), g);
        std::fclose(g);
    }
}

// the host-side sequence of hzk_g16_load, with the key arrays read to host memory
static std::string load_host(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return "cannot open";
    keyfile::Reader r{f};
    keyfile::Header h{};
    const char* bad = keyfile::read_header(r, h);
    HostCsr m[3];
    if (!bad) bad = keyfile::read_r1cs(r, h, m);
    if (!bad) {
        const uint64_t words = (h.num_vars * (64 + 64 + 128) + keyfile::l_count(h) * 64 + keyfile::h_count(h) * 64) / 8;
        U arrays(words);
        r.get(arrays.data(), words * 8);
        host::G1x p1[3];
        host::G2x p2[3];
        for (auto& p : p1) p = keyfile::get_point<host::G1x>(r);
        for (auto& p : p2) p = keyfile::get_point<host::G2x>(r);
        std::vector<host::G1x> ic(h.num_public + 1);
        for (auto& p : ic) p = keyfile::get_point<host::G1x>(r);
        bad = keyfile::read_trailer(r);
    }
    std::fclose(f);
    return bad ? bad : "ok";
}

static void check_keyfiles(const std::string& dir) {
    const std::string p = dir + "/key.bin";
    KeyImage k = make_image(6);
    write_image(p, k);
    CHECK(load_host(p) == "ok", "valid key file loads");
    auto expect = [&](const KeyImage& img, const std::string& want, bool fix = true, long trunc = -1, bool extra = false) {
        write_image(p, img, fix, trunc, extra);
        std::string got = load_host(p);
        if (got != want) std::fprintf(stderr, "  key file: got '%s', want '%s'\n", got.c_str(), want.c_str());
        CHECK(got == want, "key-file defect detected");
    };
    { KeyImage t = k; t.h.magic[0] = 'X'; expect(t, "not a proving-key file"); }
    { KeyImage t = k; t.h.version = 2; expect(t, "unsupported version"); }
    { KeyImage t = k; t.h.log_m = 30; expect(t, "inconsistent header"); }
    { KeyImage t = k; t.h.M = 48; expect(t, "inconsistent header"); }
    { KeyImage t = k; t.h.num_public = t.h.num_vars; expect(t, "inconsistent header"); }
    { KeyImage t = k; t.h.nnz[1] = (uint64_t)1 << 40; expect(t, "inconsistent header"); }
    { KeyImage t = k; t.h.num_gates = t.h.M + 1; expect(t, "inconsistent header"); }
    expect(k, "truncated header", true, 40);
    expect(k, "truncated", true, 88 + 16);
    expect(k, "trailing bytes", true, -1, true);
    expect(k, "checksum mismatch", false);
    {   // a flipped bit inside the key arrays of a valid file
        write_image(p, k);
        FILE* f = std::fopen(p.c_str(), "r+b");
        long off = (long)(sizeof(keyfile::Header) + 3 * k.h.M * 8 + (k.h.nnz[0] + k.h.nnz[1] + k.h.nnz[2]) * 40 + 100);
        std::fseek(f, off, SEEK_SET);
        int c = std::fgetc(f);
        std::fseek(f, off, SEEK_SET);
        std::fputc(c ^ 0x10, f);
        std::fclose(f);
        std::string got = load_host(p);
        CHECK(got == "checksum mismatch", "flipped key bit detected");
    }
    { KeyImage t = k; t.m[0].row_ptr[3] = t.h.nnz[0] + 1; expect(t, "corrupt row pointers"); }
    { KeyImage t = k; std::swap(t.m[1].row_ptr[4], t.m[1].row_ptr[5]); if (t.m[1].row_ptr[4] != t.m[1].row_ptr[5]) expect(t, "corrupt row pointers"); }
    { KeyImage t = k; t.m[1].col[2] = t.h.num_vars; expect(t, "column index out of range"); }
    { KeyImage t = k; t.m[2].row_ptr[1] = t.m[2].row_ptr[0]; expect(t, "a constraint row without a single output entry"); }
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: host_sanitize <scratch dir>\n");
        return 2;
    }
    check_fields();
    check_points_and_synth();
    check_host_pool();
    check_fr29();
    check_oracle_msm_ntt_pairing();
    check_groth16(6);
    check_keyfiles(argv[1]);
    if (failures) {
        std::fprintf(stderr, "host_sanitize: %d failure(s)\n", failures);
        return 1;
    }
    std::printf("host_sanitize ok\n");
    return 0;
}
