// Host-only part of the proving-key file format (hzk_g16_save / hzk_g16_load):
// header, checksum, R1CS section and their validation. No HIP: the loader in
// keyfile.cpp and the sanitizer build of the host code (tests/cpp) share it.
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
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include "host_ec.hpp"

namespace hzk {

// one R1CS matrix on the host: row_ptr has M entries and no sentinel (the
// last row ends at nnz), the reference's CSR (xpu-spmat.cuh:20-25)
struct HostCsr {
    std::vector<uint64_t> row_ptr, col;
    std::vector<host::Fr> val;
};

namespace keyfile {

inline constexpr char MAGIC[8] = {'H', 'Z', 'K', 'G', '1', '6', 0, 1};
inline constexpr uint32_t VERSION = 1;

struct Sum {
    uint64_t h = 0xcbf29ce484222325ULL;
    unsigned char pend[8];
    int np = 0;
    void word(uint64_t w) { h = (h ^ w) * 0x100000001b3ULL + ((h << 29) | (h >> 35)); }
    void add(const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        while (n && np) {
            pend[np++] = *b++;
            --n;
            if (np == 8) {
                uint64_t w;
                std::memcpy(&w, pend, 8);
                word(w);
                np = 0;
            }
        }
        for (; n >= 8; n -= 8, b += 8) {
            uint64_t w;
            std::memcpy(&w, b, 8);
            word(w);
        }
        while (n--) pend[np++] = *b++;
    }
    uint64_t done() {
        if (np) {
            std::memset(pend + np, 0, 8 - np);
            uint64_t w;
            std::memcpy(&w, pend, 8);
            word(w);
            np = 0;
        }
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

// element counts of the five key arrays: A1, B1, B2 (num_vars each), L1, H1
inline uint64_t l_count(const Header& h) { return h.num_vars - h.num_public - 1; }
inline uint64_t h_count(const Header& h) { return h.M - 1; }

// nullptr when the header's fields are consistent, else the reason
inline const char* header_problem(const Header& h) {
    if (std::memcmp(h.magic, MAGIC, 8) != 0) return "not a proving-key file";
    if (h.version != VERSION) return "unsupported version";
    if (h.log_m < 4 || h.log_m > 24 || h.M != ((uint64_t)1 << h.log_m) || h.kind > 1 ||
        h.num_public + 1 >= h.num_vars || h.num_gates > h.M || h.num_inputs >= h.num_vars ||
        h.num_vars > ((uint64_t)1 << 32) || h.nnz[0] > ((uint64_t)1 << 36) || h.nnz[1] > ((uint64_t)1 << 36) ||
        h.nnz[2] > ((uint64_t)1 << 36))
        return "inconsistent header";
    return nullptr;
}

// the whole file size, which follows from a consistent header
inline uint64_t expected_size(const Header& h) {
    uint64_t want = sizeof(Header) + 8;
    for (int i = 0; i < 3; ++i) want += h.M * 8 + h.nnz[i] * 40;
    want += h.num_vars * (64 + 64 + 128) + l_count(h) * 64 + h_count(h) * 64 + 3 * 64 + 3 * 128 +
            (h.num_public + 1) * 64;
    return want;
}

// Reads the header from r and checks it and the file's size (before anything
// is allocated from it); r is left at the R1CS section.
inline const char* read_header(Reader& r, Header& h) {
    r.get(&h, sizeof h);
    if (!r.ok) return "truncated header";
    if (const char* bad = header_problem(h)) return bad;
    long size = -1;
    if (std::fseek(r.f, 0, SEEK_END) == 0) size = std::ftell(r.f);
    if (size < 0 || std::fseek(r.f, (long)sizeof(Header), SEEK_SET) != 0) return "cannot seek";
    const uint64_t want = expected_size(h);
    if ((uint64_t)size != want) return (uint64_t)size < want ? "truncated" : "trailing bytes";
    return nullptr;
}

// Reads the three CSR matrices and checks them: row pointers monotone and
// within nnz, columns within the variables, and each constraint row r <
// num_gates holding a single C entry (the witness generator evaluates row r
// into that entry's variable).
inline const char* read_r1cs(Reader& r, const Header& h, HostCsr m[3]) {
    for (int i = 0; i < 3; ++i) {
        HostCsr& c = m[i];
        c.row_ptr.resize(h.M);
        c.col.resize(h.nnz[i]);
        c.val.resize(h.nnz[i]);
        r.get(c.row_ptr.data(), h.M * 8);
        r.get(c.col.data(), h.nnz[i] * 8);
        r.get(c.val.data(), h.nnz[i] * 32);
        if (!r.ok) return "truncated";
        for (size_t j = 0; j < h.M; ++j)
            if (c.row_ptr[j] > h.nnz[i] || (j && c.row_ptr[j] < c.row_ptr[j - 1])) return "corrupt row pointers";
        for (size_t j = 0; j < h.nnz[i]; ++j)
            if (c.col[j] >= h.num_vars) return "column index out of range";
        for (size_t row = 0; i == 2 && row < h.num_gates; ++row) {
            const uint64_t e = row + 1 < h.M ? c.row_ptr[row + 1] : h.nnz[i];
            if (e != c.row_ptr[row] + 1) return "a constraint row without a single output entry";
        }
    }
    return nullptr;
}

// After every section has gone through r: the trailing checksum matches and
// nothing follows it.
inline const char* read_trailer(Reader& r) {
    if (!r.ok) return "truncated";
    const uint64_t want = r.sum.done();
    uint64_t got = 0;
    if (std::fread(&got, 1, 8, r.f) != 8) return "truncated";
    if (got != want) return "checksum mismatch";
    char extra;
    if (std::fread(&extra, 1, 1, r.f) == 1) return "trailing bytes";
    return nullptr;
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

}  // namespace keyfile
}  // namespace hzk
