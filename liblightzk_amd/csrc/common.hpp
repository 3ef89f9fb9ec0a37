// Library context: device, stream, scratch arena, per-kernel timing.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>
#include "../../include/hzk.h"

struct hzk_graph_entry {
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;
};

// autotuned MSM parameters for a size class (log2 n); c = 0: use the heuristic
struct hzk_tune_entry {
    int c = 0;
    uint32_t K = 0;
    uint32_t g = 0;
};

struct hzk_ctx {
    int device = 0;
    int num_cus = 256;
    size_t lds_per_block = 65536;   // device limit, read at creation
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    // scratch arena: grown on demand, reused across calls, freed at destroy
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    uint64_t scratch_generation = 0;   // bumped on every reallocation (invalidates captured graphs)
    // small device buffer for batch descriptors of graph-captured work (ntt.hip)
    void* desc = nullptr;
    size_t desc_bytes = 0;
    uint64_t desc_generation = 0;
    // captured device-stage graphs keyed by batch shape
    std::map<std::string, hzk_graph_entry> graphs;
    // pinned host staging slots, used round-robin: a slot is handed out again
    // only after the copies that read it (fenced by its event) have run, so
    // back-to-back batched calls never overwrite descriptors still in flight
    struct Stage { void* p = nullptr; size_t bytes = 0; hipEvent_t done = nullptr; bool fenced = false; };
    static constexpr int STAGE_SLOTS = 4;
    Stage stage[STAGE_SLOTS];
    int stage_next = 0;
    // timing: when enabled every kernel launch is bracketed by events
    bool timing = false;
    struct Rec { hipEvent_t a, b; std::string name; };
    std::vector<Rec> pending;
    std::map<std::string, std::pair<double, long>> totals;   // name -> (ms, launches)
    std::vector<hipEvent_t> event_pool;
    // fault injection for tests of the error paths (host-side only)
    long fail_alloc_after = -1;
    // fixed-base comb tables (device, built on first use)
    void* comb_g1 = nullptr;
    void* comb_g2 = nullptr;
    // affine copy of a reference-layout base table (layout.hip)
    void* conv = nullptr;
    size_t conv_bytes = 0;
    // last MSM: entries, segments, largest bucket, window bits, windows
    uint32_t last_msm[5] = {0, 0, 0, 0, 0};
    // NTT twiddle / coset tables for the last transform size (ntt.hip)
    void* ntt_cache = nullptr;
    // MSM plans [regime][G1, G2][log2 n]; regime 0: an MSM that has the GPU
    // to itself (latency), 1: an instance of a batch wide enough to fill the
    // GPU by itself (throughput)
    hzk_tune_entry tune[2][2][32];
    int force_regime = -1;        // -1: by batch width; 0 / 1: forced (autotuner)
    // second context (own stream and scratch) for work overlapped with this one's
    hzk_ctx* aux = nullptr;
    // autotune overrides (0 = automatic)
    int force_window = 0;
    int force_seg = 0;
};

namespace hzk {

int set_error(hzk_ctx* ctx, int code, const char* fmt, ...);

#define HZK_HIP(ctx, expr)                                                            \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
            return ::hzk::set_error((ctx), HZK_ERR_HIP, "%s failed: %s (%s:%d)", #expr, \
                                    hipGetErrorString(_e), __FILE__, __LINE__);       \
    } while (0)

#define HZK_CHECK(ctx, cond, code, ...)                                   \
    do {                                                                  \
        if (!(cond)) return ::hzk::set_error((ctx), (code), __VA_ARGS__); \
    } while (0)

// Scratch arena bump allocator for one operation (256-byte aligned slices).
struct Arena {
    char* base;
    size_t size, used = 0;
    void* take(size_t bytes) {
        size_t off = (used + 255) & ~size_t(255);
        used = off + bytes;
        return base + off;
    }
};
// Make sure ctx->scratch holds at least `bytes`; returns 0 or an error code.
int ensure_scratch(hzk_ctx* ctx, size_t bytes);
// maps the static scratch of hzk_msm_g1_noalloc for a device (msm_g1.hip)
int msm_noalloc_init(int device);
int ctx_aux(hzk_ctx* ctx, hzk_ctx** out);
void timing_merge(hzk_ctx* ctx, hzk_ctx* from);
// Pinned host staging of at least `bytes`: *buf is writable now; call
// stage_fence(ctx, slot) after enqueueing the copies that read it.
int stage_acquire(hzk_ctx* ctx, size_t bytes, char** buf, int* slot);
int stage_fence(hzk_ctx* ctx, int slot);
void ntt_release(hzk_ctx* ctx);
int spmv3(hzk_ctx* ctx, const hzk_csr* A, const hzk_csr* B, const hzk_csr* C, const void* z, void* Az, void* Bz, void* Cz);
int pointwise_abc(hzk_ctx* ctx, size_t n, void* A, const void* B, const void* C, const void* scale);
// dst[k] = src[idx[k]], k < n, for elements of elem_bytes (a multiple of 16) on ctx's stream
int gather_rows(hzk_ctx* ctx, const void* src, const uint32_t* idx, size_t n, size_t elem_bytes, void* dst);
// dst[k] = z[idx[k]] * r over Fr (Montgomery; r: 32 bytes on the host), k < n
int gather_scale_fr(hzk_ctx* ctx, const void* z, const uint32_t* idx, size_t n, const void* r_host, void* dst);
// Copy `bytes` of host descriptors into ctx->desc (grown stream-ordered when
// short) through pinned staging, on ctx's stream; *dev points at them there.
// The next upload reuses the buffer: enqueue the kernels that read these first.
int upload_desc(hzk_ctx* ctx, const void* host, size_t bytes, const void** dev);

// The per-request vector stages of several proofs, one launch each (spmv.hip).
struct SpmvJob {   // Az, Bz, Cz = A z, B z, C z for one request
    const hzk_csr* m[3];
    const void* z;
    void* out[3];
};
int spmv3_batch(hzk_ctx* ctx, const SpmvJob* jobs, size_t count);
struct AbcJob {    // A <- (A o B - C) * scale over n elements
    void* A;
    const void* B;
    const void* C;
    const void* scale;   // device Fr or null
    size_t n;
};
int pointwise_abc_batch(hzk_ctx* ctx, const AbcJob* jobs, size_t count);
struct GatherJob { // dst[k] = src[idx[k]] (idx null: src[k]) for n elements of elem_bytes; times r (Fr) when scale
    const void* src;
    const uint32_t* idx;
    void* dst;
    size_t n, elem_bytes;
    const void* r_host;  // 32 bytes, Montgomery, or null; only with elem_bytes = 32
};
int gather_batch(hzk_ctx* ctx, const GatherJob* jobs, size_t count);
// in-place standard NTT (inverse: w^-1 and 1/n) and coset NTT over Fr (ntt.hip)
int ntt_standard_raw(hzk_ctx* ctx, void* x, size_t deg, int inverse);
struct NttJob {
    void* x;
    size_t deg;
    int kind;   // 0 NTT, 1 inverse NTT, 2 coset NTT, 3 coset inverse NTT (natural order);
                // 4..6 the H pipeline's permutation-free forms (ntt.hip, ntt_batch)
};
int ntt_batch(hzk_ctx* ctx, const NttJob* jobs, size_t count, int capture);
// reference base tables (layout.hip): level 0 as affine in ctx->conv
int ref_bases_affine(hzk_ctx* ctx, int group, const void* ref, size_t n, const void** out);
int coset_ntt_raw(hzk_ctx* ctx, void* x, size_t deg, int inverse);

// Kernel timing helpers (no-ops unless ctx->timing).
struct TimedLaunch {
    hzk_ctx* ctx;
    hipEvent_t b = nullptr;
    TimedLaunch(hzk_ctx* c, const char* name);
    ~TimedLaunch();
};

inline unsigned blocks_for(size_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace hzk
