// Proof-serving loop with continuous batching (no counterpart in the
// reference, which proves once per process: test/groth16.cu:3-8).
//
//   * Several circuits are set up once; their proving keys stay in HBM.
//   * Clients submit (circuit, witness, r|s) from any thread; each request is
//     stamped on arrival and on completion (latency = completion - arrival).
//   * Several pipelines (a worker thread with its own two streams and MSM
//     scratch) share one admission queue, so one pipeline's host phases
//     (Horner combine, assembly) and latency-bound kernels overlap another's
//     accumulation. Each pipeline runs steps. A step admits pending requests in arrival order
//     while their device footprint fits the memory budget (a request that can
//     never fit fails; one that does not fit now waits), then runs two stages
//     concurrently on two streams:
//       stage 1  newly admitted requests: upload z, SpMV, and the H pipeline
//                with every request's transforms in shared batched launches;
//       stage 2  requests whose H finished in the previous step: their three
//                MSMs (G1: the compacted A query and the merged L | H | B1 MSM
//                of C; G2: the compacted B query) as one batched G1 launch
//                sequence (2 per request) and one G2 sequence, then the host
//                assembly.
//     Requests that arrive while others are on the GPU join at the next step,
//     so the device always holds the H work of the newest requests next to the
//     MSM work of the previous ones.
//   * mode 1 (sequential) runs one request at a time through both stages with
//     no overlap: the plain prover driven through the same queue, the baseline
//     the batched loop is compared with.
// A failing stage fails only the requests in it; keys stay resident and the
// loop continues. Every proof is a deterministic function of (circuit, z, r, s),
// so a request's bytes are the same alone and in any mix.
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>
#include "g16.hpp"
#include "host_pool.hpp"

using namespace hzk;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id;
    int circuit;
    std::vector<uint8_t> z;   // witness, pageable (when no pinned buffer was free)
    void* zp = nullptr;       // witness in a pooled pinned buffer
    size_t zbytes = 0;
    uint8_t rs[64];
    uint8_t proof[256];
    int status = 0;   // 0 pending, 1 done, < 0 failed
    double t_arrive = 0, t_done = 0;
    size_t footprint = 0;
    void *dz = nullptr, *da = nullptr, *db = nullptr, *dc = nullptr;
};

// device buffers of one request (z, three evaluation vectors)
struct BufSet {
    void *z, *a, *b, *c;
};

}  // namespace

struct Pipeline {
    int index = 0;
    hzk_ctx* ctx_h = nullptr;   // stage 1 stream
    hzk_ctx* ctx_m = nullptr;   // stage 2 stream (and its auxiliary context for G2)
    std::thread th;
};

struct hzk_server {
    int device = 0;
    int mode = 0;
    int max_batch = 64;
    size_t budget = 0, in_use = 0;
    bool auto_budget = false;   // budget from free memory at start (after the keys are resident)
    std::vector<std::unique_ptr<Pipeline>> pipes;
    std::vector<hzk_g16*> circuits;
    // finished requests' buffers, kept per circuit for reuse (hipFree would
    // synchronise the device and stall the other stage)
    std::vector<std::vector<BufSet>> free_bufs;
    size_t cached_bytes = 0;
    // pinned host buffers for witnesses, pooled per circuit: a copy from
    // pageable memory is staged synchronously by the runtime and would hold
    // the worker thread (and the GPU behind it) for the whole transfer
    std::vector<std::vector<void*>> pinned_free;
    size_t pinned_bytes = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t next_id = 1;
    bool running = false, stop = false;
    long inflight = 0;
    std::string last_error;
    // statistics
    long steps = 0, completed = 0, failed = 0, max_inflight = 0, batched_sum = 0, msm_stages = 0;
};

namespace {

// z, a (M elements), b and c (vec_b / vec_c elements: the H stage's B z and
// C z, then the scalars of the A, merged C and B2 MSMs)
size_t bufset_bytes(const hzk_g16* g) { return (g->num_vars + g->M + g->vec_b + g->vec_c) * 32; }

void release(BufSet& b) {
    for (void* p : {b.z, b.a, b.b, b.c})
        if (p) (void)hipFree(p);
    b = {nullptr, nullptr, nullptr, nullptr};
}

// Buffers for r: reused from its circuit's free list, else allocated after
// trimming other circuits' cached buffers so that requests in flight plus the
// cache stay within the budget.
int alloc_request(hzk_server* s, Request* r) {
    hzk_g16* g = s->circuits[r->circuit];
    const size_t bytes = bufset_bytes(g);
    BufSet b{nullptr, nullptr, nullptr, nullptr};
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto& fl = s->free_bufs[r->circuit];
        if (!fl.empty()) {
            b = fl.back();
            fl.pop_back();
            s->cached_bytes -= bytes;
        }
    }
    if (!b.z) {
        for (;;) {
            BufSet victim{nullptr, nullptr, nullptr, nullptr};
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->in_use + s->cached_bytes <= s->budget) break;
                for (size_t c = 0; c < s->free_bufs.size() && !victim.z; ++c)
                    if (!s->free_bufs[c].empty()) {
                        victim = s->free_bufs[c].back();
                        s->free_bufs[c].pop_back();
                        s->cached_bytes -= bufset_bytes(s->circuits[c]);
                    }
            }
            if (!victim.z) break;
            release(victim);
        }
        size_t zb = g->num_vars * 32, vb = g->M * 32;
        if (hipMalloc(&b.z, zb) != hipSuccess || hipMalloc(&b.a, vb) != hipSuccess ||
            hipMalloc(&b.b, g->vec_b * 32) != hipSuccess || hipMalloc(&b.c, g->vec_c * 32) != hipSuccess) {
            (void)hipGetLastError();
            release(b);
            return HZK_ERR_NOMEM;
        }
    }
    r->dz = b.z;
    r->da = b.a;
    r->db = b.b;
    r->dc = b.c;
    return 0;
}

constexpr size_t PINNED_CAP = (size_t)2 << 30;   // pinned witness bytes at most

// witness of r into a pinned buffer of its circuit's pool (a new one while
// the pool is under PINNED_CAP), else into pageable memory
void stage_witness(hzk_server* s, Request* r, const void* src) {
    void* p = nullptr;
    bool grow = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto& fl = s->pinned_free[r->circuit];
        if (!fl.empty()) {
            p = fl.back();
            fl.pop_back();
        } else if (s->pinned_bytes + r->zbytes <= PINNED_CAP) {
            s->pinned_bytes += r->zbytes;
            grow = true;
        }
    }
    if (grow) {
        (void)hipSetDevice(s->device);
        if (hipHostMalloc(&p, r->zbytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            std::lock_guard<std::mutex> lk(s->mu);
            s->pinned_bytes -= r->zbytes;
        }
    }
    if (p) {
        std::memcpy(p, src, r->zbytes);
        r->zp = p;
    } else {
        r->z.assign((const uint8_t*)src, (const uint8_t*)src + r->zbytes);
    }
}

// the pinned witness buffer back to its pool; only once its upload is done
void unstage_witness_locked(hzk_server* s, Request* r) {
    if (!r->zp) return;
    s->pinned_free[r->circuit].push_back(r->zp);
    r->zp = nullptr;
}
void unstage_witness(hzk_server* s, std::vector<Request*>& rs) {
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request* r : rs) unstage_witness_locked(s, r);
}

// back to the circuit's free list (caller holds no lock)
void free_request(hzk_server* s, Request* r) {
    if (!r->dz) return;
    BufSet b{r->dz, r->da, r->db, r->dc};
    r->dz = r->da = r->db = r->dc = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    s->free_bufs[r->circuit].push_back(b);
    s->cached_bytes += bufset_bytes(s->circuits[r->circuit]);
}

void finish(hzk_server* s, std::vector<Request*>& rs, int status, const hzk_ctx* err_ctx = nullptr) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (status < 0 && !rs.empty()) s->last_error = err_ctx ? hzk_last_error(err_ctx) : "device buffer allocation failed";
    double t = now_s();
    for (Request* r : rs) {
        r->status = status;
        r->t_done = t;
        s->in_use -= r->footprint;
        --s->inflight;
        if (status == 1) ++s->completed; else ++s->failed;
    }
    if (status == 1) {
        s->batched_sum += (long)rs.size();
        ++s->msm_stages;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();   // memory freed: waiting pipelines may admit
}

// stage 1: upload z and enqueue SpMV + H pipeline for the batch (ctx_h stream)
int stage_h(hzk_server* s, hzk_ctx* ctx, std::vector<Request*>& rs) {
    if (rs.empty()) return 0;
    std::vector<G16Req> reqs;
    for (Request* r : rs) {
        const void* src = r->zp ? r->zp : (const void*)r->z.data();
        HZK_HIP(ctx, hipMemcpyAsync(r->dz, src, r->zbytes, hipMemcpyHostToDevice, ctx->stream));
        reqs.push_back({s->circuits[r->circuit], r->dz, r->da, r->db, r->dc, r->rs, {}});
    }
    return g16_h_batch(ctx, reqs.data(), reqs.size());
}

// stage 2: batched MSMs and host assembly (ctx_m stream; returns when done)
int stage_msm(hzk_server* s, hzk_ctx* ctx, std::vector<Request*>& rs) {
    if (rs.empty()) return 0;
    std::vector<G16Req> reqs;
    for (Request* r : rs) reqs.push_back({s->circuits[r->circuit], r->dz, r->da, r->db, r->dc, r->rs, {}});
    // the key-and-(r, s) terms of every proof on the host while the device runs the MSMs
    std::vector<G16Fixed> fx(rs.size());
    std::thread fixed([&] { parallel_for((long)rs.size(), [&](long i) { fx[i] = g16_fixed_terms(reqs[i].g, rs[i]->rs); }); });
    int rc = g16_msm_batch(ctx, reqs.data(), reqs.size(), 0);
    fixed.join();
    if (rc) return rc;
    parallel_for((long)rs.size(), [&](long i) { g16_assemble(reqs[i].g, reqs[i].res, rs[i]->rs, fx[i], rs[i]->proof); });
    return 0;
}

void worker_loop(hzk_server* s, Pipeline* p) {
    (void)hipSetDevice(s->device);
    std::vector<Request*> stage1, stage2;
    int mode = 0;
    for (;;) {
        std::vector<Request*> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // in sequential mode only the first pipeline proves
            auto may_take = [&] { return !s->pending.empty() && (s->mode == 0 || p->index == 0); };
            s->cv_work.wait(lk, [&] { return s->stop || may_take() || !stage1.empty(); });
            if (s->stop && !may_take() && stage1.empty()) break;
            mode = s->mode;
            size_t cap = mode == 1 ? 1 : (size_t)s->max_batch;
            while (may_take() && admitted.size() < cap) {
                Request* r = s->pending.front();
                if (r->footprint > s->budget) {   // can never fit: fail it
                    s->pending.pop_front();
                    unstage_witness_locked(s, r);
                    r->status = HZK_ERR_NOMEM;
                    r->t_done = now_s();
                    ++s->failed;
                    s->cv_done.notify_all();
                    continue;
                }
                if (s->in_use + r->footprint > s->budget) break;   // wait for memory
                s->pending.pop_front();
                s->in_use += r->footprint;
                ++s->inflight;
                admitted.push_back(r);
            }
            if (s->inflight > s->max_inflight) s->max_inflight = s->inflight;
            if (admitted.empty() && stage1.empty()) {   // memory held by another pipeline: wait for a finish
                s->cv_work.wait(lk);
                continue;
            }
            ++s->steps;
        }
        // device buffers for the admitted requests
        std::vector<Request*> ok;
        for (Request* r : admitted) {
            if (alloc_request(s, r)) {
                std::vector<Request*> one{r};
                unstage_witness(s, one);
                finish(s, one, HZK_ERR_NOMEM);
            } else {
                ok.push_back(r);
            }
        }
        if (mode == 1) {
            // plain: one request through both stages, nothing overlapped
            const hzk_ctx* ec = p->ctx_h;
            int rc = stage_h(s, p->ctx_h, ok);
            const bool synced = hipStreamSynchronize(p->ctx_h->stream) == hipSuccess;
            if (!rc && !synced) rc = HZK_ERR_HIP;
            unstage_witness(s, ok);   // uploads are done (or the stream failed)
            if (!rc) {
                ec = p->ctx_m;
                rc = stage_msm(s, p->ctx_m, ok);
            }
            for (Request* r : ok) free_request(s, r);
            finish(s, ok, rc ? rc : 1, ec);
            continue;
        }
        stage2.swap(stage1);
        stage1.swap(ok);
        // stage 1 enqueued first so its kernels run while stage 2's MSMs do
        int rc1 = stage_h(s, p->ctx_h, stage1);
        int rc2 = stage_msm(s, p->ctx_m, stage2);
        if (hipStreamSynchronize(p->ctx_h->stream) != hipSuccess && !rc1) rc1 = HZK_ERR_HIP;
        unstage_witness(s, stage1);
        for (Request* r : stage2) free_request(s, r);
        finish(s, stage2, rc2 ? rc2 : 1, p->ctx_m);
        stage2.clear();
        if (rc1) {
            for (Request* r : stage1) free_request(s, r);
            finish(s, stage1, rc1, p->ctx_h);
            stage1.clear();
        }
    }
}

void destroy_pipes(hzk_server* s) {
    for (auto& p : s->pipes) {
        hzk_ctx_destroy(p->ctx_h);
        hzk_ctx_destroy(p->ctx_m);
        p->ctx_h = p->ctx_m = nullptr;
    }
}

int add_pipe(hzk_server* s) {
    auto p = std::make_unique<Pipeline>();
    p->index = (int)s->pipes.size();
    int rc = hzk_ctx_create(s->device, nullptr, &p->ctx_h);
    if (!rc) rc = hzk_ctx_create(s->device, nullptr, &p->ctx_m);
    if (rc) {
        hzk_ctx_destroy(p->ctx_h);
        return rc;
    }
    s->pipes.push_back(std::move(p));
    return 0;
}

}  // namespace

extern "C" {

int hzk_server_create(int device, size_t mem_budget, int max_batch, int mode, hzk_server** out) {
    if (!out || max_batch < 1 || (mode != 0 && mode != 1)) return HZK_ERR_INVALID;
    *out = nullptr;
    hzk_server* s = new hzk_server();
    s->device = device;
    s->mode = mode;
    s->max_batch = max_batch;
    // three pipelines by default: with two, the GPU idled 4-13 ms at a time
    // while both were in their host phases (served mix: 554 vs 566-576 proofs/s)
    int rc = 0;
    for (int i = 0; i < 3 && !rc; ++i) rc = add_pipe(s);
    if (rc) {
        destroy_pipes(s);
        delete s;
        return rc;
    }
    s->budget = mem_budget;
    s->auto_budget = mem_budget == 0;
    *out = s;
    return HZK_OK;
}

int hzk_server_set_pipelines(hzk_server* s, int n) {
    if (!s || n < 1 || n > 8 || s->running) return HZK_ERR_INVALID;
    while ((int)s->pipes.size() > n) {
        hzk_ctx_destroy(s->pipes.back()->ctx_h);
        hzk_ctx_destroy(s->pipes.back()->ctx_m);
        s->pipes.pop_back();
    }
    while ((int)s->pipes.size() < n) {
        int rc = add_pipe(s);
        if (rc) return rc;
    }
    return HZK_OK;
}

const char* hzk_server_last_error(const hzk_server* s) { return s ? s->last_error.c_str() : "null server"; }

int hzk_server_add_circuit(hzk_server* s, uint64_t seed, int log_m, int num_public, int* circuit_id) {
    if (!s || !circuit_id || s->running) return HZK_ERR_INVALID;
    hzk_g16* g = nullptr;
    hzk_ctx* ctx = s->pipes[0]->ctx_m;
    int rc = hzk_g16_setup(ctx, seed, log_m, num_public, &g);
    if (rc) {
        s->last_error = hzk_last_error(ctx);
        return rc;
    }
    s->circuits.push_back(g);
    s->free_bufs.emplace_back();
    s->pinned_free.emplace_back();
    *circuit_id = (int)s->circuits.size() - 1;
    return HZK_OK;
}

// Add a circuit from a proving-key file written by hzk_g16_save.
int hzk_server_add_circuit_file(hzk_server* s, const char* path, int* circuit_id) {
    if (!s || !circuit_id || s->running) return HZK_ERR_INVALID;
    hzk_g16* g = nullptr;
    hzk_ctx* ctx = s->pipes[0]->ctx_m;
    int rc = hzk_g16_load(ctx, path, &g);
    if (rc) {
        s->last_error = hzk_last_error(ctx);
        return rc;
    }
    s->circuits.push_back(g);
    s->free_bufs.emplace_back();
    s->pinned_free.emplace_back();
    *circuit_id = (int)s->circuits.size() - 1;
    return HZK_OK;
}

// Fixed-base level tables for a circuit's key (hzk_g16_set_levels), before start.
int hzk_server_set_levels(hzk_server* s, int circuit_id, int levels) {
    if (!s || s->running || circuit_id < 0 || circuit_id >= (int)s->circuits.size()) return HZK_ERR_INVALID;
    hzk_ctx* ctx = s->pipes[0]->ctx_m;
    int rc = hzk_g16_set_levels(ctx, s->circuits[circuit_id], levels);
    if (rc) s->last_error = hzk_last_error(ctx);
    return rc;
}

int hzk_server_circuit(hzk_server* s, int circuit_id, hzk_g16** out) {
    if (!s || !out || circuit_id < 0 || circuit_id >= (int)s->circuits.size()) return HZK_ERR_INVALID;
    *out = s->circuits[circuit_id];
    return HZK_OK;
}

int hzk_server_footprint(hzk_server* s, int circuit_id, size_t* bytes) {
    if (!s || !bytes || circuit_id < 0 || circuit_id >= (int)s->circuits.size()) return HZK_ERR_INVALID;
    *bytes = g16_request_footprint(s->circuits[circuit_id]);
    return HZK_OK;
}

int hzk_server_set_mode(hzk_server* s, int mode) {
    if (!s || (mode != 0 && mode != 1)) return HZK_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->pending.empty() || s->in_use != 0) {
        s->last_error = "set_mode: requests in flight";
        return HZK_ERR_INVALID;
    }
    s->mode = mode;
    s->cv_work.notify_all();
    return HZK_OK;
}

int hzk_server_start(hzk_server* s) {
    if (!s || s->running) return HZK_ERR_INVALID;
    if (s->auto_budget) {
        size_t fr = 0, tot = 0;
        (void)hipSetDevice(s->device);
        (void)hipMemGetInfo(&fr, &tot);
        s->budget = fr / 2;
    }
    // pinned witness buffers up front (an equal share of the cap per circuit,
    // at most what the pipelines can hold in flight): allocating pinned memory
    // while requests flow serializes with the device work of the pipelines
    (void)hipSetDevice(s->device);
    const size_t ncirc = s->circuits.size(), in_flight = 2 * (size_t)s->max_batch * s->pipes.size();
    for (size_t c = 0; c < ncirc; ++c) {
        const size_t zbytes = s->circuits[c]->num_vars * 32;
        const size_t want = std::min(in_flight, PINNED_CAP / ncirc / zbytes);
        while (s->pinned_free[c].size() < want && s->pinned_bytes + zbytes <= PINNED_CAP) {
            void* p = nullptr;
            if (hipHostMalloc(&p, zbytes, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                break;
            }
            s->pinned_free[c].push_back(p);
            s->pinned_bytes += zbytes;
        }
    }
    s->running = true;
    for (auto& p : s->pipes) p->th = std::thread(worker_loop, s, p.get());
    return HZK_OK;
}

int hzk_server_submit(hzk_server* s, int circuit_id, const void* z_host, const void* rs_host, uint64_t* request_id) {
    if (!s || !z_host || !rs_host || !request_id || circuit_id < 0 || circuit_id >= (int)s->circuits.size())
        return HZK_ERR_INVALID;
    auto r = std::make_unique<Request>();
    hzk_g16* g = s->circuits[circuit_id];
    r->circuit = circuit_id;
    r->zbytes = g->num_vars * 32;
    stage_witness(s, r.get(), z_host);
    std::memcpy(r->rs, rs_host, 64);
    r->footprint = g16_request_footprint(g);
    std::lock_guard<std::mutex> lk(s->mu);
    r->t_arrive = now_s();
    r->id = s->next_id++;
    *request_id = r->id;
    s->pending.push_back(r.get());
    s->all[r->id] = std::move(r);
    s->cv_work.notify_all();
    return HZK_OK;
}

// status: 1 done (proof written), 0 still pending after timeout, < 0 failed.
// A finished request is forgotten once its result has been returned.
int hzk_server_wait(hzk_server* s, uint64_t request_id, double timeout_s, void* proof_out, double* latency_s) {
    if (!s) return HZK_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return HZK_ERR_INVALID;
    Request* r = it->second.get();
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->status != 0; });
    int st = r->status;
    if (st == 1 && proof_out) std::memcpy(proof_out, r->proof, 256);
    if (latency_s) *latency_s = st != 0 ? r->t_done - r->t_arrive : 0.0;
    if (st != 0) {
        unstage_witness_locked(s, r);
        s->all.erase(it);
    }
    return st;
}

// {steps, completed, failed, max in flight, mean requests per MSM stage, budget bytes, bytes admitted}
int hzk_server_stats(hzk_server* s, double* out7) {
    if (!s || !out7) return HZK_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    out7[0] = (double)s->steps;
    out7[1] = (double)s->completed;
    out7[2] = (double)s->failed;
    out7[3] = (double)s->max_inflight;
    out7[4] = s->msm_stages ? (double)s->batched_sum / (double)s->msm_stages : 0.0;
    out7[5] = (double)s->budget;
    out7[6] = (double)s->in_use;
    return HZK_OK;
}

int hzk_server_inject_failure(hzk_server* s, long after_n) {
    if (!s) return HZK_ERR_INVALID;
    return hzk_inject_alloc_failure(s->pipes[0]->ctx_m, after_n);
}

void hzk_server_destroy(hzk_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto& p : s->pipes)
        if (p->th.joinable()) p->th.join();
    for (auto& kv : s->all) {
        free_request(s, kv.second.get());
        if (kv.second->zp) (void)hipHostFree(kv.second->zp);
    }
    for (auto& fl : s->free_bufs)
        for (BufSet& b : fl) release(b);
    for (auto& fl : s->pinned_free)
        for (void* p : fl) (void)hipHostFree(p);
    for (hzk_g16* g : s->circuits) hzk_g16_free(g);
    destroy_pipes(s);
    delete s;
}

}  // extern "C"
