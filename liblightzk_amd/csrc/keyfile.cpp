// On-disk proving key and R1CS (hzk_g16_save / hzk_g16_load).
//
// The reference regenerates random keys in every run (include/protocol/
// groth16.cuh:84-122) and has no key format. This file format lets a prover
// or a proof server restart without repeating Setup (the QAP evaluation and
// the fixed-base generation of every key point).
//
// The byte layout, checksum and the host-side checks live in
// keyfile_format.hpp; this file moves the key arrays between file and HBM.
#include <cstdio>
#include <cstring>
#include <vector>
#include "common.hpp"
#include "g16.hpp"
#include "keyfile_format.hpp"

using namespace hzk;

namespace {

using keyfile::Header;
using keyfile::Reader;
using keyfile::Writer;
const size_t STAGE = (size_t)64 << 20;   // host staging chunk for key arrays

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
    g->own(d, bytes);
    *out = d;
    for (size_t off = 0; off < bytes && r.ok; off += STAGE) {
        const size_t n = std::min(STAGE, bytes - off);
        r.get(stage.data(), n);
        if (r.ok) HZK_HIP(ctx, hipMemcpy((char*)d + off, stage.data(), n, hipMemcpyHostToDevice));
    }
    return HZK_OK;
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
    std::memcpy(h.magic, keyfile::MAGIC, 8);
    h.version = keyfile::VERSION;
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
    keyfile::put_point(w, g->alpha1);
    keyfile::put_point(w, g->beta1);
    keyfile::put_point(w, g->delta1);
    keyfile::put_point(w, g->beta2);
    keyfile::put_point(w, g->gamma2);
    keyfile::put_point(w, g->delta2);
    for (const host::G1x& p : g->IC) keyfile::put_point(w, p);
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
    const char* bad = keyfile::read_header(r, h);
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
    bad = keyfile::read_r1cs(r, h, g->m);
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
        g->alpha1 = keyfile::get_point<host::G1x>(r);
        g->beta1 = keyfile::get_point<host::G1x>(r);
        g->delta1 = keyfile::get_point<host::G1x>(r);
        g->beta2 = keyfile::get_point<host::G2x>(r);
        g->gamma2 = keyfile::get_point<host::G2x>(r);
        g->delta2 = keyfile::get_point<host::G2x>(r);
        g->IC.resize(g->num_public + 1);
        for (host::G1x& p : g->IC) p = keyfile::get_point<host::G1x>(r);
    }
    if (!rc && !bad) bad = keyfile::read_trailer(r);
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
                    g->own(d, parts[k].second);
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
