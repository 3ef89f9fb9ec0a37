// MSM autotuner: per size class (floor(log2 n)) and group, the window bits c,
// the accumulation chunk length K and the bucket-reduction run length g are
// chosen by timing candidate plans on this GPU. No counterpart in the
// reference, which hard-wires c = 12 and its launch geometry
// (include/operator/msm.cuh:121, :225, :245).
//
// Two tables, one per regime: plans for an MSM that has the GPU to itself
// (single MSMs and narrow batches), timed alone, and plans for the instances
// of a wide batch (the throughput regime, see msm_layout), timed as a batch of
// B = 2^26 / (16 n) instances of the class.
//
// Search: coordinate descent from the current plan (tuned entry, built-in
// table or heuristic): first c over [c0 - 3, c0 + 2] (with the heuristic K and
// g of each c, or the current K and g in the batch regime), then K over
// {K/4 .. 4K}, then g over powers of two (up to 32, or 64 in the batch
// regime). Each candidate is the best of several timed runs after one
// warm-up, on seeded uniform scalars and bases k_i * G generated on the
// device. The winner is written to the context's table of that regime.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>
#include "common.hpp"
#include "msm.hpp"

using namespace hzk;

namespace {

// one MSM (batch = 1) or a batch of `batch` instances of n points each, the
// i-th over scalars offset by i*step (distinct digit streams, shared bases)
double time_msm(hzk_ctx* ctx, int group, const void* s, const void* b, size_t n, uint32_t batch, size_t step,
                void* dest, int reps, int* rc) {
    std::vector<MsmJob> jobs(batch);
    for (uint32_t i = 0; i < batch; ++i)
        jobs[i] = MsmJob{(const char*)s + 32 * step * i, b, n, (char*)dest + 256 * (size_t)i, 0, 0};
    auto run = [&]() {
        if (batch == 1) return group == 1 ? msm_g1(ctx, s, b, n, 0, dest) : msm_g2(ctx, s, b, n, 0, dest);
        return group == 1 ? msm_batch_g1(ctx, jobs.data(), batch, 0, 0) : msm_batch_g2(ctx, jobs.data(), batch, 0, 0);
    };
    *rc = run();
    if (*rc) return 0;
    double best = 1e30;
    for (int r = 0; r < reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        *rc = run();
        if (*rc) return 0;
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) { *rc = HZK_ERR_HIP; return 0; }
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el < best) best = el;
    }
    return best;
}

int tune_set(hzk_ctx* ctx, int regime, int group, int log_n, int c, uint32_t K, uint32_t g) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, (group == 1 || group == 2) && log_n >= 0 && log_n < 32, HZK_ERR_INVALID, "tune_set: group %d log_n %d",
              group, log_n);
    HZK_CHECK(ctx, c == 0 || (c >= MSM_MIN_C && c <= MSM_MAX_C), HZK_ERR_INVALID, "tune_set: c = %d", c);
    HZK_CHECK(ctx, K == 0 || (K % 4 == 0 && K <= (1u << 16)), HZK_ERR_INVALID, "tune_set: K = %u (multiple of 4)", K);
    HZK_CHECK(ctx, g == 0 || ((g & (g - 1)) == 0 && g <= 4096 && (c == 0 || g <= (1u << (c - 1)))), HZK_ERR_INVALID,
              "tune_set: g = %u (power of two, at most 2^(c-1))", g);
    hzk_tune_entry& t = ctx->tune[regime][group - 1][log_n];
    t.c = c;
    t.K = c ? K : 0;
    t.g = c ? g : 0;
    return HZK_OK;
}

int tune_get(hzk_ctx* ctx, int regime, int group, int log_n, int* c, uint32_t* K, uint32_t* g) {
    if (!ctx || !c || !K || !g) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, (group == 1 || group == 2) && log_n >= 0 && log_n < 32, HZK_ERR_INVALID, "tune_get: group %d log_n %d",
              group, log_n);
    MsmPlan p = msm_plan_ctx(ctx, group - 1, (size_t)1 << log_n, 0, regime);
    *c = p.c;
    *K = p.K;
    *g = p.g;
    return HZK_OK;
}

// Batch regime: a class is timed as a batch of B = 2^26 / (16 n) instances
// (2 to 256), about the width at which a batch takes the throughput plans.
uint32_t tune_batch_width(int log_n) {
    uint64_t b = ((uint64_t)1 << 26) / (((uint64_t)16) << log_n);
    return (uint32_t)(b < 2 ? 2 : (b > 256 ? 256 : b));
}

int autotune(hzk_ctx* ctx, int regime, int group, int log_min, int log_max, double budget_ms) {
    if (!ctx) return HZK_ERR_INVALID;
    HZK_CHECK(ctx, (group == 1 || group == 2) && log_min >= 1 && log_min <= log_max && log_max <= 26, HZK_ERR_INVALID,
              "autotune: group %d, classes %d..%d", group, log_min, log_max);
    const size_t nmax = (size_t)1 << log_max;
    const size_t pt = group == 1 ? 64 : 128;
    // scalars: the batch regime offsets instance i by i*step scalars
    size_t smax = nmax;
    if (regime == MSM_REGIME_BATCH)
        for (int L = log_min; L <= log_max; ++L) {
            const size_t n = (size_t)1 << L, step = n / 64 + 1;
            smax = std::max(smax, n + step * (tune_batch_width(L) - 1));
        }
    const uint32_t maxb = regime == MSM_REGIME_BATCH ? tune_batch_width(log_min) : 1;
    // inputs: seeded scalars (stream 0) and base multipliers (stream 1) of seed 77
    std::vector<uint64_t> host(4 * smax);
    void *ds = nullptr, *dk = nullptr, *db = nullptr, *dd = nullptr;
    auto cleanup = [&]() {
        for (void* p : {ds, dk, db, dd})
            if (p) (void)hipFree(p);
    };
    int rc = hzk_gen_field(77, 1, nmax, 0, host.data());
    if (!rc) rc = hzk_device_alloc(ctx, 32 * nmax, &dk);
    if (!rc) rc = hzk_memcpy_h2d(ctx, dk, host.data(), 32 * nmax);
    if (!rc) rc = hzk_device_alloc(ctx, pt * nmax, &db);
    if (!rc) rc = group == 1 ? hzk_fixbase_g1(ctx, dk, nmax, db) : hzk_fixbase_g2(ctx, dk, nmax, db);
    if (!rc) rc = hzk_gen_field(77, 0, smax, 0, host.data());
    if (!rc) rc = hzk_device_alloc(ctx, 32 * smax, &ds);
    if (!rc) rc = hzk_memcpy_h2d(ctx, ds, host.data(), 32 * smax);
    if (!rc) rc = hzk_device_alloc(ctx, 256 * (size_t)maxb, &dd);
    if (rc) { cleanup(); return rc; }
    const int saved_w = ctx->force_window, saved_s = ctx->force_seg, saved_r = ctx->force_regime;
    ctx->force_window = ctx->force_seg = 0;
    ctx->force_regime = regime;
    int cmax = MSM_MAX_C;
    while (cmax > MSM_MIN_C && ((size_t)4 << (cmax - 1)) > ctx->lds_per_block) --cmax;

    for (int L = log_min; L <= log_max && !rc; ++L) {
        const size_t n = (size_t)1 << L;
        const uint32_t B = regime == MSM_REGIME_BATCH ? tune_batch_width(L) : 1;
        const size_t step = n / 64 + 1;
        hzk_tune_entry& slot = ctx->tune[regime][group - 1][L];
        const int reps = L <= 16 ? 6 : (L <= 20 ? 4 : 2);
        auto t_class = std::chrono::steady_clock::now();
        auto over = [&]() {
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_class).count() > budget_ms;
        };
        MsmPlan base = msm_plan_ctx(ctx, group - 1, n, 0, regime);
        hzk_tune_entry best{base.c, base.K, base.g};
        slot = best;
        double tbest = time_msm(ctx, group, ds, db, n, B, step, dd, reps, &rc);
        auto try_entry = [&](hzk_tune_entry e) {
            if (rc || over()) return;
            slot = e;
            double t = time_msm(ctx, group, ds, db, n, B, step, dd, reps, &rc);
            if (!rc && t < tbest) { tbest = t; best = e; }
            slot = best;
        };
        // 1. window bits (K, g from the heuristic for that c, or the current
        //    K and g in the batch regime, where lanes are plentiful)
        for (int c = base.c - 3; c <= base.c + 2; ++c) {
            if (c < MSM_MIN_C || c > cmax || c == base.c) continue;
            MsmPlan q = msm_plan(n, c, 0, ctx->lds_per_block);
            if (regime == MSM_REGIME_BATCH) try_entry({c, base.K, std::min<uint32_t>(base.g, 1u << (c - 1))});
            else try_entry({c, q.K, q.g});
        }
        // 2. chunk length
        const hzk_tune_entry afterc = best;
        for (uint32_t K : {afterc.K / 4, afterc.K / 2, afterc.K * 2, afterc.K * 4}) {
            if (K < 4 || K % 4 || K > 4096) continue;
            try_entry({afterc.c, K, afterc.g});
        }
        // 3. bucket-reduction run length
        const hzk_tune_entry afterk = best;
        for (uint32_t g = 1; g <= 64; g <<= 1) {
            if (g == afterk.g || g > (1u << (afterk.c - 1))) continue;
            if (regime == MSM_REGIME_SINGLE && g > 32) continue;
            try_entry({afterk.c, afterk.K, g});
        }
        slot = best;
    }
    ctx->force_window = saved_w;
    ctx->force_seg = saved_s;
    ctx->force_regime = saved_r;
    cleanup();
    return rc;
}

}  // namespace

extern "C" {

int hzk_msm_tune_set(hzk_ctx* ctx, int group, int log_n, int c, uint32_t K, uint32_t g) {
    return tune_set(ctx, MSM_REGIME_SINGLE, group, log_n, c, K, g);
}
int hzk_msm_tune_get(hzk_ctx* ctx, int group, int log_n, int* c, uint32_t* K, uint32_t* g) {
    return tune_get(ctx, MSM_REGIME_SINGLE, group, log_n, c, K, g);
}
int hzk_msm_autotune(hzk_ctx* ctx, int group, int log_min, int log_max, double budget_ms) {
    return autotune(ctx, MSM_REGIME_SINGLE, group, log_min, log_max, budget_ms);
}
int hzk_msm_batch_tune_set(hzk_ctx* ctx, int group, int log_n, int c, uint32_t K, uint32_t g) {
    return tune_set(ctx, MSM_REGIME_BATCH, group, log_n, c, K, g);
}
int hzk_msm_batch_tune_get(hzk_ctx* ctx, int group, int log_n, int* c, uint32_t* K, uint32_t* g) {
    return tune_get(ctx, MSM_REGIME_BATCH, group, log_n, c, K, g);
}
int hzk_msm_batch_autotune(hzk_ctx* ctx, int group, int log_min, int log_max, double budget_ms) {
    return autotune(ctx, MSM_REGIME_BATCH, group, log_min, log_max, budget_ms);
}

}  // extern "C"
