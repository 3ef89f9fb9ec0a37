// C ABI: context, buffers, timing, the device field-arithmetic unit entry,
// and the MSM entry points (declarations and reference citations in
// include/hzk.h).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>
#include "common.hpp"
#include "curve.hpp"
#include "msm.hpp"

namespace hzk {

int set_error(hzk_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    return code;
}

int ensure_scratch(hzk_ctx* ctx, size_t bytes) {
    if (ctx->fail_alloc_after >= 0) {
        if (ctx->fail_alloc_after == 0) {
            ctx->fail_alloc_after = -1;
            return set_error(ctx, HZK_ERR_NOMEM, "scratch allocation of %zu bytes failed (injected)", bytes);
        }
        --ctx->fail_alloc_after;
    }
    if (ctx->scratch_bytes >= bytes) return 0;
    ++ctx->scratch_generation;
    // Stream-ordered: the old arena is released after the work already
    // enqueued on this stream, and neither call waits for the device (a
    // hipFree would stall every stream of the process, e.g. a server's other
    // pipelines). Growth at least doubles, so a varying batch mix settles
    // after a few steps.
    const size_t old = ctx->scratch_bytes;
    if (ctx->scratch) {
        HZK_HIP(ctx, hipFreeAsync(ctx->scratch, ctx->stream));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = std::max(bytes + bytes / 8, 2 * old);
    hipError_t e = hipMallocAsync(&ctx->scratch, want, ctx->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipMallocAsync(&ctx->scratch, bytes, ctx->stream);
        want = bytes;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->scratch = nullptr;
        return set_error(ctx, HZK_ERR_NOMEM, "scratch allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    ctx->scratch_bytes = want;
    return 0;
}

int upload_desc(hzk_ctx* ctx, const void* host, size_t bytes, const void** dev) {
    if (ctx->desc_bytes < bytes) {   // stream-ordered like the scratch arena: no device-wide stall
        if (ctx->desc) HZK_HIP(ctx, hipFreeAsync(ctx->desc, ctx->stream));
        ctx->desc = nullptr;
        ctx->desc_bytes = 0;
        const size_t want = std::max<size_t>({bytes * 2, (size_t)1 << 16});
        HZK_HIP(ctx, hipMallocAsync(&ctx->desc, want, ctx->stream));
        ctx->desc_bytes = want;
        ++ctx->desc_generation;
    }
    char* pin;
    int slot;
    int rc = stage_acquire(ctx, bytes, &pin, &slot);
    if (rc) return rc;
    std::memcpy(pin, host, bytes);
    HZK_HIP(ctx, hipMemcpyAsync(ctx->desc, pin, bytes, hipMemcpyHostToDevice, ctx->stream));
    rc = stage_fence(ctx, slot);
    *dev = ctx->desc;
    return rc;
}

int stage_acquire(hzk_ctx* ctx, size_t bytes, char** buf, int* slot) {
    int i = ctx->stage_next;
    ctx->stage_next = (i + 1) % hzk_ctx::STAGE_SLOTS;
    hzk_ctx::Stage& st = ctx->stage[i];
    if (st.fenced) {
        HZK_HIP(ctx, hipEventSynchronize(st.done));
        st.fenced = false;
    }
    if (!st.done) HZK_HIP(ctx, hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
    if (st.bytes < bytes) {
        const size_t old = st.bytes;
        if (st.p) HZK_HIP(ctx, hipHostFree(st.p));
        st.p = nullptr;
        st.bytes = 0;
        size_t want = std::max<size_t>({(size_t)1 << 20, bytes + bytes / 4, 2 * old});   // pinned allocations are slow: grow rarely
        HZK_HIP(ctx, hipHostMalloc(&st.p, want, hipHostMallocDefault));
        st.bytes = want;
    }
    *buf = (char*)st.p;
    *slot = i;
    return 0;
}

int stage_fence(hzk_ctx* ctx, int slot) {
    hzk_ctx::Stage& st = ctx->stage[slot];
    HZK_HIP(ctx, hipEventRecord(st.done, ctx->stream));
    st.fenced = true;
    return 0;
}

static hipEvent_t take_event(hzk_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static void drain_timing(hzk_ctx* ctx) {
    for (auto& r : ctx->pending) {
        float ms = 0;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto& t = ctx->totals[r.name];
            t.first += ms;
            t.second += 1;
        }
        ctx->event_pool.push_back(r.a);
        ctx->event_pool.push_back(r.b);
    }
    ctx->pending.clear();
}

// the auxiliary context (second stream and scratch) used to overlap a G2 batch
// with a G1 batch; inherits the timing switch
int ctx_aux(hzk_ctx* ctx, hzk_ctx** out) {
    if (!ctx->aux) {
        int rc = hzk_ctx_create(ctx->device, nullptr, &ctx->aux);
        if (rc) return set_error(ctx, rc, "auxiliary context: creation failed");
    }
    std::memcpy(ctx->aux->tune, ctx->tune, sizeof ctx->tune);
    ctx->aux->timing = ctx->timing;
    ctx->aux->fail_alloc_after = -1;
    *out = ctx->aux;
    return 0;
}

// fold the auxiliary context's kernel times into ctx's totals
void timing_merge(hzk_ctx* ctx, hzk_ctx* from) {
    drain_timing(from);
    for (auto& kv : from->totals) {
        auto& t = ctx->totals[kv.first];
        t.first += kv.second.first;
        t.second += kv.second.second;
    }
    from->totals.clear();
}

TimedLaunch::TimedLaunch(hzk_ctx* c, const char* name) : ctx(c) {
    if (!ctx->timing) return;
    hipEvent_t a = take_event(ctx);
    b = take_event(ctx);
    if (!a || !b) { b = nullptr; return; }
    (void)hipEventRecord(a, ctx->stream);
    ctx->pending.push_back({a, b, name});
}
TimedLaunch::~TimedLaunch() {
    if (b) (void)hipEventRecord(b, ctx->stream);
}

// ------------------------------------------------------------------ field unit kernels
template <class F>
__global__ void k_field_op(int op, const F* a, const F* b, F* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F x = a[i], r;
    switch (op) {
        case 0: r = x * b[i]; break;
        case 1: r = x + b[i]; break;
        case 2: r = x - b[i]; break;
        case 3: r = x.neg(); break;
        case 4: r = x.sqr(); break;
        case 5: r = x.inverse(); break;
        default: r = x; break;
    }
    out[i] = r;
}
template <class F>
__global__ void k_field_mont(int op, const F* a, F* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = op == 6 ? a[i].from_mont() : a[i].to_mont();
}

}  // namespace hzk

using namespace hzk;

extern "C" {

const char* hzk_version(void) { return "hzk 0.1 gfx950"; }

int hzk_ctx_create(int device, void* stream, hzk_ctx** out) {
    if (!out) return HZK_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HZK_ERR_HIP;
    if (device < 0 || device >= ndev) return HZK_ERR_INVALID;
    hzk_ctx* ctx = new hzk_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return HZK_ERR_HIP; }
    if (msm_noalloc_init(device) != HZK_OK) { delete ctx; return HZK_ERR_HIP; }
    {   // scratch arenas come from the device's stream-ordered pool; freed ones stay cached there
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        (void)hipGetLastError();
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->lds_per_block = prop.sharedMemPerBlock;
        ctx->num_cus = prop.multiProcessorCount;
    }
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return HZK_ERR_HIP; }
        ctx->own_stream = true;
    }
    *out = ctx;
    return HZK_OK;
}

void hzk_ctx_destroy(hzk_ctx* ctx) {
    if (!ctx) return;
    hzk_ctx_destroy(ctx->aux);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drain_timing(ctx);
    for (auto& kv : ctx->graphs)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->scratch) {
        (void)hipFreeAsync(ctx->scratch, ctx->stream);
        (void)hipStreamSynchronize(ctx->stream);
    }
    ntt_release(ctx);
    if (ctx->desc) {
        (void)hipFreeAsync(ctx->desc, ctx->stream);
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (ctx->comb_g1) (void)hipFree(ctx->comb_g1);
    if (ctx->comb_g2) (void)hipFree(ctx->comb_g2);
    if (ctx->conv) (void)hipFree(ctx->conv);
    for (auto& st : ctx->stage) {
        if (st.p) (void)hipHostFree(st.p);
        if (st.done) (void)hipEventDestroy(st.done);
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* hzk_last_error(const hzk_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }
void* hzk_ctx_stream(hzk_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int hzk_sync(hzk_ctx* ctx) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HZK_OK;
}

int hzk_device_sync(hzk_ctx* ctx) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipSetDevice(ctx->device));
    HZK_HIP(ctx, hipDeviceSynchronize());
    return HZK_OK;
}

int hzk_device_alloc(hzk_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return HZK_ERR_INVALID;
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return set_error(ctx, HZK_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return HZK_OK;
}
int hzk_device_free(hzk_ctx* ctx, void* ptr) {
    if (!ctx) return HZK_ERR_INVALID;
    if (!ptr) return HZK_OK;
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HZK_HIP(ctx, hipFree(ptr));
    return HZK_OK;
}
int hzk_memcpy_h2d(hzk_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HZK_OK;
}
int hzk_memcpy_d2h(hzk_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HZK_OK;
}
int hzk_memset(hzk_ctx* ctx, void* dst, int value, size_t bytes) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
    return HZK_OK;
}
int hzk_mem_info(hzk_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipMemGetInfo(free_bytes, total_bytes));
    return HZK_OK;
}

int hzk_field_op(hzk_ctx* ctx, int field, int op, const void* a, const void* b, void* out, size_t n) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, field >= 0 && field <= 2 && op >= 0 && op <= 7, HZK_ERR_INVALID, "field_op: bad field/op");
    HZK_CHECK(ctx, a && out && (b || op >= 3), HZK_ERR_INVALID, "field_op: null pointer");
    HZK_CHECK(ctx, !(field == 2 && op >= 6), HZK_ERR_UNSUPPORTED, "field_op: Fq2 has no Montgomery conversion");
    if (n == 0) return HZK_OK;
    dim3 grid(blocks_for(n, 256)), block(256);
    TimedLaunch tl(ctx, "field_op");
    if (op >= 6) {
        if (field == 0) hipLaunchKernelGGL(k_field_mont<fq_t>, grid, block, 0, ctx->stream, op, (const fq_t*)a, (fq_t*)out, n);
        else hipLaunchKernelGGL(k_field_mont<fr_t>, grid, block, 0, ctx->stream, op, (const fr_t*)a, (fr_t*)out, n);
    } else if (field == 0) {
        hipLaunchKernelGGL(k_field_op<fq_t>, grid, block, 0, ctx->stream, op, (const fq_t*)a, (const fq_t*)b, (fq_t*)out, n);
    } else if (field == 1) {
        hipLaunchKernelGGL(k_field_op<fr_t>, grid, block, 0, ctx->stream, op, (const fr_t*)a, (const fr_t*)b, (fr_t*)out, n);
    } else {
        hipLaunchKernelGGL(k_field_op<fq2_t>, grid, block, 0, ctx->stream, op, (const fq2_t*)a, (const fq2_t*)b, (fq2_t*)out, n);
    }
    HZK_HIP(ctx, hipGetLastError());
    return HZK_OK;
}

int hzk_msm_g1(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, uint32_t base_layout, size_t win_siz,
               void* dest) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, base_layout <= HZK_BASES_AFFINE_LEVELS, HZK_ERR_UNSUPPORTED, "msm_g1: base layout %u unsupported", base_layout);
    if (base_layout == HZK_BASES_REF_LEVELS) {
        HZK_CHECK(ctx, bases && n > 0, HZK_ERR_INVALID, "msm_g1: null bases or n = 0");
        int rc = ref_bases_affine(ctx, 1, bases, n, &bases);
        if (rc) return rc;
    }
    if (base_layout == HZK_BASES_AFFINE_LEVELS) {
        HZK_CHECK(ctx, win_siz >= MSM_MIN_C && win_siz <= MSM_MAX_C, HZK_ERR_INVALID, "msm_g1: a level table needs its win_siz");   // full table: ceil(255 / c) levels
        return msm_g1(ctx, scalars, bases, n, win_siz, dest, (uint32_t)((255 + win_siz - 1) / win_siz));
    }
    return msm_g1(ctx, scalars, bases, n, win_siz, dest);
}
int hzk_msm_g2(hzk_ctx* ctx, const void* scalars, const void* bases, size_t n, uint32_t base_layout, size_t win_siz,
               void* dest) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, base_layout <= HZK_BASES_AFFINE_LEVELS, HZK_ERR_UNSUPPORTED, "msm_g2: base layout %u unsupported", base_layout);
    if (base_layout == HZK_BASES_REF_LEVELS) {
        HZK_CHECK(ctx, bases && n > 0, HZK_ERR_INVALID, "msm_g2: null bases or n = 0");
        int rc = ref_bases_affine(ctx, 2, bases, n, &bases);
        if (rc) return rc;
    }
    if (base_layout == HZK_BASES_AFFINE_LEVELS) {
        HZK_CHECK(ctx, win_siz >= MSM_MIN_C && win_siz <= MSM_MAX_C, HZK_ERR_INVALID, "msm_g2: a level table needs its win_siz");
        return msm_g2(ctx, scalars, bases, n, win_siz, dest, (uint32_t)((255 + win_siz - 1) / win_siz));
    }
    return msm_g2(ctx, scalars, bases, n, win_siz, dest);
}

static int msm_batch_entry(hzk_ctx* ctx, const hzk_msm_desc* d, size_t count, uint32_t flags, bool g2) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, d && count > 0, HZK_ERR_INVALID, "msm batch: empty");
    HZK_CHECK(ctx, (flags & ~(HZK_BATCH_HOST_DEST | HZK_BATCH_GRAPH | HZK_BATCH_DEVICE_COMBINE)) == 0, HZK_ERR_INVALID,
              "msm batch: bad flags %u", flags);
    HZK_CHECK(ctx, !((flags & HZK_BATCH_HOST_DEST) && (flags & HZK_BATCH_DEVICE_COMBINE)), HZK_ERR_INVALID,
              "msm batch: the device-side combine writes device results (no HZK_BATCH_HOST_DEST)");
    std::vector<MsmJob> jobs(count);
    for (size_t i = 0; i < count; ++i) jobs[i] = {d[i].scalars, d[i].bases, d[i].n, d[i].dest, d[i].win_siz, d[i].levels};
    int host = (flags & HZK_BATCH_HOST_DEST) ? 1 : 0, graph = (flags & HZK_BATCH_GRAPH) ? 1 : 0;
    int dev = (flags & HZK_BATCH_DEVICE_COMBINE) ? 1 : 0;
    return g2 ? msm_batch_g2(ctx, jobs.data(), count, host, graph, dev) : msm_batch_g1(ctx, jobs.data(), count, host, graph, dev);
}
int hzk_msm_g1_batch(hzk_ctx* ctx, const hzk_msm_desc* d, size_t count, uint32_t flags) {
    return msm_batch_entry(ctx, d, count, flags, false);
}
int hzk_msm_g2_batch(hzk_ctx* ctx, const hzk_msm_desc* d, size_t count, uint32_t flags) {
    return msm_batch_entry(ctx, d, count, flags, true);
}

int hzk_msm_last_stats(hzk_ctx* ctx, uint32_t* out5) {
    if (!ctx || !out5) return HZK_ERR_INVALID;
    for (int i = 0; i < 5; ++i) out5[i] = ctx->last_msm[i];
    return HZK_OK;
}

int hzk_timing_enable(hzk_ctx* ctx, int enable) {
    if (!ctx) return HZK_ERR_INVALID;
    ctx->timing = enable != 0;
    return HZK_OK;
}
int hzk_timing_reset(hzk_ctx* ctx) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drain_timing(ctx);
    ctx->totals.clear();
    return HZK_OK;
}
int hzk_timing_get(hzk_ctx* ctx, const char* name, double* ms, long* launches) {
    if (!ctx || !name) return HZK_ERR_INVALID;
    HZK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    drain_timing(ctx);
    auto it = ctx->totals.find(name);
    if (ms) *ms = it == ctx->totals.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == ctx->totals.end() ? 0 : it->second.second;
    return HZK_OK;
}
size_t hzk_timing_report(hzk_ctx* ctx, char* buf, size_t cap) {
    if (!ctx) return 0;
    (void)hipStreamSynchronize(ctx->stream);
    drain_timing(ctx);
    std::string s;
    for (auto& kv : ctx->totals) {
        char line[256];
        snprintf(line, sizeof line, "%s %.6f %ld\n", kv.first.c_str(), kv.second.first, kv.second.second);
        s += line;
    }
    if (buf && cap) {
        size_t m = s.size() < cap - 1 ? s.size() : cap - 1;
        memcpy(buf, s.data(), m);
        buf[m] = 0;
    }
    return s.size() + 1;
}

int hzk_set_tuning(hzk_ctx* ctx, int window_bits, int segment) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, window_bits == 0 || (window_bits >= MSM_MIN_C && window_bits <= MSM_MAX_C), HZK_ERR_INVALID,
              "window_bits %d out of range", window_bits);
    HZK_CHECK(ctx, segment == 0 || (segment > 0 && (segment & (segment - 1)) == 0), HZK_ERR_INVALID,
              "segment %d must be a power of two", segment);
    ctx->force_window = window_bits;
    ctx->force_seg = segment;
    return HZK_OK;
}
int hzk_set_msm_regime(hzk_ctx* ctx, int regime) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, regime >= -1 && regime <= 1, HZK_ERR_INVALID, "msm regime %d (-1 auto, 0 single, 1 wide batch)", regime);
    ctx->force_regime = regime;
    return HZK_OK;
}
int hzk_inject_alloc_failure(hzk_ctx* ctx, long after_n) {
    if (!ctx) return HZK_ERR_INVALID;
    ctx->fail_alloc_after = after_n;
    return HZK_OK;
}

}  // extern "C"
