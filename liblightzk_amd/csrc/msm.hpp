// MSM plan, batch descriptors and entry points shared by the C ABI, the
// batcher and the prover.
#pragma once
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

struct hzk_ctx;

namespace hzk {

constexpr int MSM_MIN_C = 4;
constexpr int MSM_MAX_C = 16;   // 2^(c-1) LDS counters per histogram workgroup

struct MsmPlan {
    int c;          // window bits
    int W;          // windows, ceil(255 / c)
    uint32_t NB;    // buckets per window, 2^(c-1)
    uint32_t K;     // sorted entries per accumulation lane (chunk length)
    uint32_t g;     // buckets per bucket-reduction run
    uint32_t T;     // scalar tiles per window (histogram / scatter workgroups)
    uint32_t tile;  // scalars per tile
};
// Plan for n points: autotuned entry for the size class if one is loaded,
// else the built-in heuristic; force_c / force_g override (0 = automatic).
MsmPlan msm_plan(size_t n, int force_c, int force_g, size_t lds_bytes);
// plan regimes: an MSM with the GPU to itself, or an instance of a wide
// batch: two or more instances whose digits total at least
// MSM_WIDE_BATCH_ENTRIES, or many instances (MSM_WIDE_BATCH_COUNT, G1 / G2)
// totalling MSM_WIDE_BATCH_MIN_ENTRIES (measured on served-like batches: a
// batch of 16 G1 or 16 G2 MSMs of mixed sizes already runs faster on the
// throughput plans, while the 4 G1 MSMs of one 2^18 proof do not)
constexpr int MSM_REGIME_SINGLE = 0, MSM_REGIME_BATCH = 1;
constexpr uint64_t MSM_WIDE_BATCH_ENTRIES = (uint64_t)1 << 26;
constexpr uint64_t MSM_WIDE_BATCH_MIN_ENTRIES = (uint64_t)1 << 24;
constexpr size_t MSM_WIDE_BATCH_COUNT[2] = {8, 16};
inline bool msm_wide_batch(int g2, size_t count, uint64_t entries) {
    if (count < 2) return false;
    return entries >= MSM_WIDE_BATCH_ENTRIES ||
           (count >= MSM_WIDE_BATCH_COUNT[g2 ? 1 : 0] && entries >= MSM_WIDE_BATCH_MIN_ENTRIES);
}
MsmPlan msm_plan_ctx(hzk_ctx* ctx, int g2, size_t n, int force_c, int regime = MSM_REGIME_SINGLE);
// the built-in plan of a regime (no context overrides)
MsmPlan msm_plan_builtin(int g2, size_t n, int regime, size_t lds_bytes);

// Work spaces of the batched stages; each instance owns a contiguous range
// [begin[s][i], begin[s][i+1]) of every space.
enum MsmSpace {
    SP_DIGIT = 0,   // one lane per scalar
    SP_TILEWG,      // one workgroup per (tile, window)
    SP_KEY,         // nk + 1 entries per instance (counts, offsets, bucket sums)
    SP_SCANTILE,    // scan tiles over the key space
    SP_CHUNK,       // accumulation chunks of K sorted entries (upper bound)
    SP_BRED,        // bucket-reduction lanes (window x run)
    SP_WIN,         // window-sum partials (S per window)
    SP_ENTRY,       // n * W digit / sorted-entry slots
    SP_HIST,        // T * nk per-tile histogram counters
    SP_CGROUP,      // W * 256 coarse key groups (two-pass scatter)
    SP_COUNT
};

// Device-side descriptor of one MSM instance of a batch.
struct MsmInst {
    const void* scalars;   // Fr[n], Montgomery
    const void* bases;     // affine points
    uint32_t n, c, W, NB, nk, K, g, G, T, tile;
    uint32_t S;            // window-sum workgroups per window
    uint32_t fused;        // 1: each window-sum thread reduces one run of g buckets itself (no run-sum pass)
    // Level tables: nd scalars with Wd digit windows over a table of Lg levels
    // (point t*nd + i = 2^(c t) P_i, t < Lg); each bucket set ("window" of the
    // pipeline) serves Lg consecutive digit windows, so W = ceil(Wd / Lg) and
    // n = nd * Lg virtual points. Plain bases: nd = n, Wd = W, Lg = 1.
    uint32_t nd, Wd, Lg;
    uint32_t tp;           // scatter in two passes (coarse groups, then buckets): 0 no, 1 yes
                           // (8-byte coarse entries), 2 yes (packed 4-byte coarse entries)
    // Digit window j covers bits [j*wb + min(j, wx), +wb + (j < wx)): the 255
    // bits of a signed recoding of a scalar < 2^254 split as evenly as the W
    // windows allow (widths c and c - 1), so the top window is not a sliver
    // of a few bits whose digits pile into a handful of buckets. Level tables
    // keep uniform windows of c bits (wb = c, wx = 0).
    uint32_t wb, wx;
    void* dest;            // the result (device), for the device-side combine
};

struct MsmBatchDev {
    const MsmInst* inst;      // [ninst]
    const uint64_t* begin;    // [SP_COUNT][ninst + 1]
    uint32_t ninst;
    uint32_t* stats;          // per instance {entries, chunks, largest bucket, 0}
};

// Launch the scalar-side stages (digits, LDS histograms, scans, scatter) for a
// whole batch. Buffers are indexed through the work spaces above.
struct MsmScalarBufs {
    uint32_t* digits;    // SP_ENTRY
    uint32_t* hist;      // SP_HIST
    uint32_t* count;     // SP_KEY
    uint32_t* offs;      // SP_KEY
    uint64_t* tiles;     // 2 * SP_SCANTILE
    uint32_t* sorted;    // SP_ENTRY
    uint2* coarse;       // SP_ENTRY 8-byte entries, or null when no instance uses the two-pass scatter
};
constexpr uint32_t MSM_COARSE = 256;   // coarse key groups per window
// Instances with many buckets per window and enough points that each coarse
// group holds a few thousand entries scatter in two passes; below that the
// per-group fixed cost outweighs the write coalescing.
// Two-pass scatter: alone, an instance pays for the extra pass from 2^19
// points on; in a batch whose instances scatter together (batch_entries
// digits in all), their random 4-byte writes share the L2 and the coalesced
// two-pass form wins from MSM_TP_BATCH_MIN_N points on.
constexpr uint64_t MSM_TP_BATCH_ENTRIES = (uint64_t)1 << 23;
constexpr uint32_t MSM_TP_BATCH_MIN_N = 1u << 15;
inline bool msm_two_pass(uint32_t n, uint32_t NB, uint64_t batch_entries) {
    if (NB < 2048) return false;
    return n >= (1u << 19) || (batch_entries >= MSM_TP_BATCH_ENTRIES && n >= MSM_TP_BATCH_MIN_N);
}
void msm_launch_scalar_stages(hzk_ctx* ctx, const MsmBatchDev& b, const uint64_t* host_totals, uint32_t max_nb,
                              const MsmScalarBufs& bufs);
constexpr uint32_t MSM_SCAN_TILE = 1024;
// one-time kernel attribute setup (call outside graph capture)
void msm_prepare();

// levels != 0: bases is a level table (see MsmJob)
int msm_g1(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest, uint32_t levels = 0);
int msm_g2(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, size_t win_siz, void* dest, uint32_t levels = 0);

// Batched MSM: out[i] (device or host pointer per dests_on_host) receives the
// normalised Jacobian result of instance i. capture != 0 records the device
// stages into a hipGraph on first use of this batch shape and replays it.
struct MsmJob {
    const void* scalars;
    const void* bases;
    size_t n;
    void* dest;
    uint32_t win_siz;
    uint32_t levels;   // 0 or 1: affine bases; else levels per bucket set of a level table (<= ceil(255 / win_siz))
    uint32_t form = 0; // 0 affine bases, 1 bases already converted by msm_convert_g1 / _g2 (kept by the prover)
};
// device_combine: the window combine runs as a kernel into device dests and
// the call returns without waiting (HZK_BATCH_DEVICE_COMBINE)
int msm_batch_g1(hzk_ctx* ctx, const MsmJob* jobs, size_t count, int dests_on_host, int capture, int device_combine = 0);
// G1 affine bases -> the form the G1 bucket accumulation reads (field29.hpp's
// 2^261 Montgomery coordinates, infinity marked), n points, on ctx's stream;
// dst may equal src
int msm_convert_g1(hzk_ctx* ctx, const void* src, void* dst, size_t n);
// the same for G2 affine bases (128 bytes each)
int msm_convert_g2(hzk_ctx* ctx, const void* src, void* dst, size_t n);
int msm_batch_g2(hzk_ctx* ctx, const MsmJob* jobs, size_t count, int dests_on_host, int capture, int device_combine = 0);

}  // namespace hzk
