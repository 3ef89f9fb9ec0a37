// On-disk proving key and R1CS (hzk_g16_save / hzk_g16_load).
//
// The reference regenerates random keys in every run (include/protocol/
// groth16.cuh:84-122) and has no key format. This file format lets a prover
// or a proof server restart without repeating Setup (the QAP evaluation and
// the fixed-base generation of every key point).
//
// Layout, all integers little-endian:
//   header   8 B magic "HZKG16\0\1", u32 version (1), u32 kind, u32 matmul p,
//            u32 log_m, u64 M, num_vars, num_public, num_gates, num_inputs,
//            u64 nnz of A, B, C                                  (88 B)
//   R1CS     for A, B, C: row_ptr u64[M], col u64[nnz], val Fr[nnz] (32 B)
//   key      A1 G1[num_vars], B1 G1[num_vars], B2 G2[num_vars],
//            L1 G1[num_vars - num_public - 1], H1 G1[M - 1]      (affine)
//   points   alpha1, beta1, delta1 (G1), beta2, gamma2, delta2 (G2),
//            IC G1[num_public + 1]                               (affine)
//   trailer  u64 checksum of every byte before it
// Affine points are 64 B (G1) / 128 B (G2) Montgomery x | y, infinity all
// zero, the layout of the device key arrays. The checksum folds 8-byte
// words: h = (h ^ w) * 0x100000001b3 + rotl(h, 29), seeded 0xcbf29ce484222325,
// the tail zero-padded.
#include <cstdio>
#include <cstring>
#include <vector>
#include "common.hpp"
#include "g16.hpp"

using namespace hzk;

namespace {

const char MAGIC[8] = {'H', 'Z', 'K', 'G', '1', '6', 0, 1};
const uint32_t VERSION = 1;
const size_t STAGE = (size_t)64 << 20;   // host staging chunk for key arrays

struct Sum {
    uint64_t h = 0xcbf29ce484222325ULL;
    unsigned char pend[8];
    int np = 0;
    void word(uint64_t w) { h = (h ^ w) * 0x100000001b3ULL + ((h << 29) | (h >> 35)); }
    void add(const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        while (n && np) { pend[np++] = *b++; --n; if (np == 8) { uint64_t w; std::memcpy(&w, pend, 8); word(w); np = 0; } }
        for (; n >= 8; n -= 8, b += 8) { uint64_t w; std::memcpy(&w, b, 8); word(w); }
        while (n--) pend[np++] = *b++;
    }
    uint64_t done() {
        if (np) { std::memset(pend + np, 0, 8 - np); uint64_t w; std::memcpy(&w, pend, 8); word(w); np = 0; }
        return h;
    }
};

struct Writer {
    FILE* f;
    Sum sum;
    bool ok = true;
    void put(const void* p, size_t n) {
        if (!ok || !n) return;
        sum.add(p, n);
        ok = std::fwrite(p, 1, n, f) == n;
    }
};

struct Reader {
    FILE* f;
    Sum sum;
    bool ok = true;
    void get(void* p, size_t n) {
        if (!ok || !n) return;
        ok = std::fread(p, 1, n, f) == n;
        if (ok) sum.add(p, n);
    }
};

struct Header {
    char magic[8];
    uint32_t version, kind, mm_p, log_m;
    uint64_t M, num_vars, num_public, num_gates, num_inputs, nnz[3];
};
static_assert(sizeof(Header) == 88, "key file header is 88 bytes");

// device array -> file through a host staging buffer
int put_device(hzk_ctx* ctx, Writer& w, const void* dev, size_t bytes, std::vector<char>& stage) {
    for (size_t off = 0; off < bytes; off += STAGE) {
        const size_t n = std::min(STAGE, bytes - off);
        HZK_HIP(ctx, hipMemcpy(stage.data(), (const char*)dev + off, n, hipMemcpyDeviceToHost));
        w.put(stage.data(), n);
    }
    return HZK_OK;
}

// file -> new device array (owned by g)
int get_device(hzk_ctx* ctx, Reader& r, hzk_g16* g, size_t bytes, void** out, std::vector<char>& stage) {
    void* d = nullptr;
    HZK_HIP(ctx, hipMalloc(&d, bytes ? bytes : 1));
    g->owned.push_back(d);
    *out = d;
    for (size_t off = 0; off < bytes && r.ok; off += STAGE) {
        const size_t n = std::min(STAGE, bytes - off);
        r.get(stage.data(), n);
        if (r.ok) HZK_HIP(ctx, hipMemcpy((char*)d + off, stage.data(), n, hipMemcpyHostToDevice));
    }
    return HZK_OK;
}

template <class P>
void put_point(Writer& w, const P& pt) {
    char b[2 * sizeof(decltype(P::X))];
    pt.store_affine(b);
    w.put(b, sizeof b);
}

template <class P>
P get_point(Reader& r) {
    char b[2 * sizeof(decltype(P::X))] = {};
    r.get(b, sizeof b);
    return P::load_affine(b);
}

}  // namespace

extern "C" {

int hzk_g16_save(hzk_ctx* ctx, const hzk_g16* g, const char* path) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, g && path, HZK_ERR_INVALID, "g16_save: null argument");
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FILE* f = std::fopen(path, "wb");
    HZK_CHECK(ctx, f, HZK_ERR_INVALID, "g16_save: cannot open %s for writing", path);
    Writer w{f};
    Header h{};
    std::memcpy(h.magic, MAGIC, 8);
    h.version = VERSION;
    h.kind = (uint32_t)g->kind;
    h.mm_p = (uint32_t)g->mm_p;
    h.log_m = (uint32_t)g->log_m;
    h.M = g->M;
    h.num_vars = g->num_vars;
    h.num_public = g->num_public;
    h.num_gates = g->num_gates;
    h.num_inputs = g->num_inputs;
    for (int i = 0; i < 3; ++i) h.nnz[i] = g->m[i].col.size();
    w.put(&h, sizeof h);
    for (int i = 0; i < 3; ++i) {
        w.put(g->m[i].row_ptr.data(), g->m[i].row_ptr.size() * 8);
        w.put(g->m[i].col.data(), g->m[i].col.size() * 8);
        w.put(g->m[i].val.data(), g->m[i].val.size() * 32);
    }
    std::vector<char> stage(STAGE);
    const size_t N1 = g->num_vars, Ln = N1 - g->num_public - 1, Hn = g->M - 1;
    int rc = put_device(ctx, w, g->A1, N1 * 64, stage);
    if (!rc) rc = put_device(ctx, w, g->B1, N1 * 64, stage);
    if (!rc) rc = put_device(ctx, w, g->B2, N1 * 128, stage);
    if (!rc) rc = put_device(ctx, w, g->L1, Ln * 64, stage);
    if (!rc) rc = put_device(ctx, w, g->H1, Hn * 64, stage);
    if (rc) {
        std::fclose(f);
        return rc;
    }
    put_point(w, g->alpha1);
    put_point(w, g->beta1);
    put_point(w, g->delta1);
    put_point(w, g->beta2);
    put_point(w, g->gamma2);
    put_point(w, g->delta2);
    for (const host::G1x& p : g->IC) put_point(w, p);
    uint64_t cs = w.sum.done();
    if (w.ok) w.ok = std::fwrite(&cs, 1, 8, f) == 8;
    const bool closed = std::fclose(f) == 0;
    HZK_CHECK(ctx, w.ok && closed, HZK_ERR_INVALID, "g16_save: write to %s failed", path);
    return HZK_OK;
}

int hzk_g16_load(hzk_ctx* ctx, const char* path, hzk_g16** out) {
    if (!ctx || !out) return HZK_ERR_INVALID;
    *out = nullptr;
    HZK_CHECK(ctx, path, HZK_ERR_INVALID, "g16_load: null path");
    FILE* f = std::fopen(path, "rb");
    HZK_CHECK(ctx, f, HZK_ERR_INVALID, "g16_load: cannot open %s", path);
    Reader r{f};
    Header h{};
    r.get(&h, sizeof h);
    const char* bad = nullptr;
    if (!r.ok) bad = "truncated header";
    else if (std::memcmp(h.magic, MAGIC, 8) != 0) bad = "not a proving-key file";
    else if (h.version != VERSION) bad = "unsupported version";
    else if (h.log_m < 4 || h.log_m > 24 || h.M != ((uint64_t)1 << h.log_m) || h.kind > 1 ||
             h.num_public + 1 >= h.num_vars || h.num_gates > h.M || h.num_inputs >= h.num_vars ||
             h.num_vars > ((uint64_t)1 << 32) || h.nnz[0] > ((uint64_t)1 << 36) || h.nnz[1] > ((uint64_t)1 << 36) ||
             h.nnz[2] > ((uint64_t)1 << 36))
        bad = "inconsistent header";
    if (!bad) {
        // the whole file size follows from the header: check it before allocating
        const uint64_t N1 = h.num_vars, Ln = N1 - h.num_public - 1, Hn = h.M - 1;
        uint64_t want = sizeof(Header) + 8;
        for (int i = 0; i < 3; ++i) want += h.M * 8 + h.nnz[i] * 40;
        want += N1 * (64 + 64 + 128) + Ln * 64 + Hn * 64 + 3 * 64 + 3 * 128 + (h.num_public + 1) * 64;
        long size = -1;
        if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
        if (size < 0 || std::fseek(f, (long)sizeof(Header), SEEK_SET) != 0) bad = "cannot seek";
        else if ((uint64_t)size != want) bad = (uint64_t)size < want ? "truncated" : "trailing bytes";
    }
    if (bad) {
        std::fclose(f);
        return set_error(ctx, HZK_ERR_INVALID, "g16_load: %s: %s", path, bad);
    }
    hzk_g16* g = new hzk_g16();
    g->device = ctx->device;
    g->kind = (int)h.kind;
    g->mm_p = (int)h.mm_p;
    g->log_m = (int)h.log_m;
    g->M = h.M;
    g->num_vars = h.num_vars;
    g->num_public = h.num_public;
    g->num_gates = h.num_gates;
    g->num_inputs = h.num_inputs;
    int rc = HZK_OK;
    for (int i = 0; i < 3 && r.ok; ++i) {
        HostCsr& m = g->m[i];
        m.row_ptr.resize(g->M);
        m.col.resize(h.nnz[i]);
        m.val.resize(h.nnz[i]);
        r.get(m.row_ptr.data(), g->M * 8);
        r.get(m.col.data(), h.nnz[i] * 8);
        r.get(m.val.data(), h.nnz[i] * 32);
        if (!r.ok) break;
        // CSR sanity: monotone row pointers within nnz, columns within the variables
        for (size_t j = 0; j < g->M && !bad; ++j)
            if (m.row_ptr[j] > h.nnz[i] || (j && m.row_ptr[j] < m.row_ptr[j - 1])) bad = "corrupt row pointers";
        for (size_t j = 0; j < h.nnz[i] && !bad; ++j)
            if (m.col[j] >= g->num_vars) bad = "column index out of range";
        // the witness generator evaluates row r < num_gates into the variable
        // of its single C entry
        for (size_t r = 0; i == 2 && r < g->num_gates && !bad; ++r) {
            const uint64_t e = r + 1 < g->M ? m.row_ptr[r + 1] : h.nnz[i];
            if (e != m.row_ptr[r] + 1) bad = "a constraint row without a single output entry";
        }
        if (bad) break;
    }
    std::vector<char> stage(STAGE);
    const size_t N1 = g->num_vars, Ln = N1 - g->num_public - 1, Hn = g->M - 1;
    if (!bad && r.ok) {
        rc = get_device(ctx, r, g, N1 * 64, &g->A1, stage);
        if (!rc) rc = get_device(ctx, r, g, N1 * 64, &g->B1, stage);
        if (!rc) rc = get_device(ctx, r, g, N1 * 128, &g->B2, stage);
        if (!rc) rc = get_device(ctx, r, g, Ln * 64, &g->L1, stage);
        if (!rc) rc = get_device(ctx, r, g, Hn * 64, &g->H1, stage);
    }
    if (!rc && !bad && r.ok) {
        g->alpha1 = get_point<host::G1x>(r);
        g->beta1 = get_point<host::G1x>(r);
        g->delta1 = get_point<host::G1x>(r);
        g->beta2 = get_point<host::G2x>(r);
        g->gamma2 = get_point<host::G2x>(r);
        g->delta2 = get_point<host::G2x>(r);
        g->IC.resize(g->num_public + 1);
        for (host::G1x& p : g->IC) p = get_point<host::G1x>(r);
    }
    uint64_t want = r.sum.done(), got = 0;
    if (!rc && !bad && r.ok && std::fread(&got, 1, 8, f) != 8) bad = "truncated";
    else if (!rc && !bad && !r.ok) bad = "truncated";
    else if (!rc && !bad && got != want) bad = "checksum mismatch";
    char extra;
    if (!rc && !bad && std::fread(&extra, 1, 1, f) == 1) bad = "trailing bytes";
    std::fclose(f);
    if (!rc && bad) rc = set_error(ctx, HZK_ERR_INVALID, "g16_load: %s: %s", path, bad);
    if (!rc) {
        for (int i = 0; i < 3 && !rc; ++i) {
            void* d = nullptr;
            const HostCsr& m = g->m[i];
            const std::pair<const void*, size_t> parts[3] = {
                {m.val.data(), m.val.size() * 32}, {m.col.data(), m.col.size() * 8}, {m.row_ptr.data(), m.row_ptr.size() * 8}};
            void** dst[3] = {&g->d_val[i], (void**)&g->d_col[i], (void**)&g->d_rp[i]};
            for (int k = 0; k < 3 && !rc; ++k) {
                hipError_t e = hipMalloc(&d, parts[k].second ? parts[k].second : 1);
                if (e == hipSuccess) {
                    g->owned.push_back(d);
                    *dst[k] = d;
                    if (parts[k].second) e = hipMemcpy(d, parts[k].first, parts[k].second, hipMemcpyHostToDevice);
                }
                if (e != hipSuccess) rc = set_error(ctx, HZK_ERR_HIP, "g16_load: %s", hipGetErrorString(e));
            }
        }
    }
    if (!rc) rc = g16_sparse_key(ctx, g);
    if (!rc) rc = g16_prover_buffers(ctx, g);
    if (rc) {
        hzk_g16_free(g);
        return rc;
    }
    *out = g;
    return HZK_OK;
}

}  // extern "C"
