#!/usr/bin/env python3
"""Benchmark: G1 MSM throughput at n = 2^24 on one MI355X.

Workload (seeded, synthetic): scalars uniform in Fr (stream 0 of seed 3),
bases P_i = k_i * G1 with k_i uniform in Fr (stream 1), generated on the GPU
by hzk_fixbase_g1 and resident in HBM before timing. One step = one full
Pippenger MSM (hzk_msm_g1: recode, LDS counting sort, bucket accumulation,
merge, bucket reduction, window combine, normalisation) over all 2^24 points.

Prints ONE JSON line. `value` is Mpoints/s (n / seconds per step).
`roofline` is for the dominant kernel (msm_accumulate): the algorithmic
32x32-bit integer multiplies of the MSM, W_MSM(n) = min_c ceil(255/c)
(11 n + 16 2^c) * 128 (signed-digit Pippenger: 11 Fq products per point and
window, 16 per bucket of the running-sum reduction, 128 32x32 products per
Fq product; independent of the plan this build runs), divided by the
kernel's average launch time from HIP events on the context's stream,
against the full-rate VALU integer-multiply peak of the MI355X (256 CUs x 64
lanes x 2.4 GHz = 3.93e13/s); the measured v_mad_u64_u32 issue rate is
reported beside it. The NTT and plain-proof legs carry roofline blocks of
the same form (W_NTT(n) = (n/2) log2 n * 128; a proof's W = its five MSMs'
W_MSM, G2 at 3 Fq products per Fq2 product, plus seven W_NTT(M)).
`cpu_baseline` times the oracle's naive MSM (per-thread s_i * P_i
accumulators, test/msm-check.cu:46-55 semantics) on a bounded sample of the
same inputs on this host's cores, and the oracle's Groth16 prover on the
m = 2^18 circuit of the plain-proof leg.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G1 MSM Mpoints/s @2^24; Groth16 proofs/s + p50/p99 latency, served vs plain"
INTMUL_PER_FQ = 128          # 8x8 product + 8x8 reduction, 32x32->64 each
FQ_PER_MADD = 11             # madd-2007-bl, the algorithmic count used for MSM work
HBM_PEAK = 8.0e12            # MI355X published HBM3E bandwidth, bytes/s
INTMUL_PEAK = 256 * 64 * 2.4e9   # MI355X: 256 CUs x 64 VALU lanes x 2.4 GHz, one 32x32 multiply per lane per clock


def w_msm(n, group=1):
    """Algorithmic int-mul count of an MSM: min_c ceil(255/c) (11 n + 16 2^c) M, M = M_Fq (G1) or 3 M_Fq (G2)."""
    best = None
    for c in range(1, 25):
        w = math.ceil(255 / c) * (11 * n + 16 * 2 ** c) * INTMUL_PER_FQ
        best = w if best is None or w < best else best
    return best * (3 if group == 2 else 1)


def w_ntt(n):
    """Algorithmic int-mul count of one n-point NTT: (n/2) log2 n M_Fr."""
    return (n // 2) * int(math.log2(n)) * INTMUL_PER_FQ


def w_proof(num_vars, num_public, M):
    """A Groth16 proof's algorithmic int-mul count: G1 MSMs over A, B1 (num_vars
    each), L (num_vars - num_public - 1) and H (M - 1), the G2 MSM over B
    (num_vars), and seven M-point NTTs (3 inverse, 3 coset, 1 coset inverse)."""
    return (2 * w_msm(num_vars) + w_msm(num_vars - num_public - 1) + w_msm(M - 1) + w_msm(num_vars, 2) + 7 * w_ntt(M))


def intmul_roofline(work, seconds, kernel, extra=None):
    achieved = work / seconds if seconds > 0 else 0.0
    r = {"bound": "valu-intmul", "kernel": kernel, "achieved": achieved / 1e12, "peak": INTMUL_PEAK / 1e12,
         "unit": "T int-mul/s (32x32->64)", "frac": achieved / INTMUL_PEAK, "work_per_launch": work,
         "peak_source": "256 CUs x 64 lanes x 2.4 GHz (MI355X published CU count and peak clock)"}
    if extra:
        r.update(extra)
    return r


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ctx, scalars_host, bases_dev, budget_s):
    """Oracle naive MSM on a bounded sample of the same inputs (the only oracle use here)."""
    from oracle import orc
    nthreads = orc.lib().orc_num_threads()
    n0 = 1 << 14
    t = time.perf_counter()
    orc.msm_g1(scalars_host[: 4 * n0].copy(), bases_dev.download(np.uint64, 8 * n0), 0)
    per_pt = (time.perf_counter() - t) / n0
    n_s = int(min(1 << 20, max(n0, budget_s / max(per_pt, 1e-9))))
    s = scalars_host[: 4 * n_s].copy()
    b = bases_dev.download(np.uint64, 8 * n_s)
    t = time.perf_counter()
    orc.msm_g1(s, b, 0)
    el = time.perf_counter() - t
    return {"value": n_s / el / 1e6, "unit": "Mpoints/s", "cores": nthreads, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"first {n_s} points of the same workload, naive per-thread double-and-add MSM "
                      f"(test/msm-check.cu:46-55 semantics), {el:.1f} s; the cost is linear in n, so the "
                      f"2^24 rate equals this one"}


def cpu_prover_baseline(ctx, log_m=18, proofs=2):
    """Part of the cpu_baseline leg: the oracle's honest Groth16 prover (CPU
    SpMV, radix-2 NTTs, CPU Pippenger MSMs, phase-6 assembly; the phase order
    of include/protocol/groth16.cuh:154-222) on a bounded circuit, next to the
    GPU plain prover on the same circuit and witnesses (proof bytes compared)."""
    import liblightzk_amd as z
    from oracle import orc
    o = orc.Groth16(seed=2, log_m=log_m, num_public=8)
    o.key_points(0)                       # key points built before the clock
    g = z.Groth16(ctx, seed=2, log_m=log_m, num_public=8)
    wits = [(o.witness(300 + i), o.rs(300 + i)) for i in range(proofs)]
    t = time.perf_counter()
    cpu = [o.prove(zz, rs) for zz, rs in wits]
    cpu_s = (time.perf_counter() - t) / proofs
    g.prove(*wits[0])
    t = time.perf_counter()
    gpu = [g.prove(zz, rs) for zz, rs in wits]
    gpu_s = (time.perf_counter() - t) / proofs
    g.close()
    return {"value": 1.0 / cpu_s, "unit": "proofs/s", "cores": orc.lib().orc_num_threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{proofs} proofs of the m=2^{log_m} multiplication-gate circuit (seed 2, the plain-proof "
                      f"leg's circuit; CPU SpMV, radix-2 NTTs, CPU Pippenger MSMs; key points built before the clock)",
            "gpu_plain_same_circuit_proofs_per_s": 1.0 / gpu_s,
            "proofs_identical": bool(all(np.array_equal(a, b) for a, b in zip(cpu, gpu)))}


def groth16_plain(ctx, log_m, proofs, levels=1):
    """Plain one-proof-at-a-time GPU prover on the synthetic circuit (seed 2)."""
    import liblightzk_amd as z
    t = time.perf_counter()
    g = z.Groth16(ctx, seed=2, log_m=log_m, num_public=8)
    if levels > 1:
        g.set_levels(levels)
    setup_s = time.perf_counter() - t
    wits = [(g.witness(100 + i), g.rs(100 + i)) for i in range(proofs + 1)]
    g.prove(*wits[0])   # warm-up (tables, scratch)
    lat = []
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    for zz, rs in wits[1:]:
        a = time.perf_counter()
        g.prove(zz, rs)
        lat.append(time.perf_counter() - a)
    el = time.perf_counter() - t0
    rep = ctx.timing_report()
    ctx.timing(False)
    info = g.info
    g.close()
    lat.sort()
    work = w_proof(info["num_vars"], info["num_public"], info["M"])
    return {"config": f"m=2^{log_m} multiplication-gate circuit, {g.info['num_vars']} variables",
            "roofline": intmul_roofline(work, el / proofs, "whole proof (wall clock per proof)"),
            "proofs": proofs, "proofs_per_s": proofs / el, "p50_ms": lat[len(lat) // 2] * 1e3,
            "p99_ms": lat[min(len(lat) - 1, int(len(lat) * 0.99))] * 1e3, "setup_s": setup_s, "key_levels": levels,
            "kernels_ms_per_proof": {k: v[0] / proofs for k, v in rep.items()}}


def groth16_reference_presets(ctx, ps=(39, 63, 100), proofs=5):
    """The reference's own statements (input_scale MatMul_p_39/63/100,
    include/protocol/groth16.cuh:11-46: the MatMul R1CS of
    xpu-spmat.cuh:33-65) proved one at a time by the plain prover, each with a
    satisfying seeded witness."""
    import liblightzk_amd as z
    out = {}
    for p in ps:
        t = time.perf_counter()
        g = z.Groth16.matmul(ctx, 2, p)
        setup_s = time.perf_counter() - t
        wits = [(g.witness(200 + i), g.rs(200 + i)) for i in range(proofs + 1)]
        g.prove(*wits[0])
        lat = []
        t0 = time.perf_counter()
        for zz, rs in wits[1:]:
            a = time.perf_counter()
            g.prove(zz, rs)
            lat.append(time.perf_counter() - a)
        el = time.perf_counter() - t0
        info = g.info
        g.close()
        out[f"MatMul_p_{p}"] = {"M": info["M"], "k": info["num_public"], "N": info["num_vars"] - 1,
                                "proofs": proofs, "proofs_per_s": proofs / el, "p50_ms": _pct(lat, 0.5) * 1e3,
                                "p99_ms": _pct(lat, 0.99) * 1e3, "setup_s": setup_s}
    return out


def served_msm_batches(ctx, reps=3, seed=4):
    """The MSM stage of 32 served requests of the 40/30/20/10 mix (13 x 2^12,
    10 x 2^16, 6 x 2^18, 3 x 2^20 constraints, n = 2^k) launched alone, with
    the prover's MSM shapes: one G1 batch of 2 MSMs per request (the compacted
    A query, 2n/3 points, and the merged C MSM, 8n/3 points: L, H and the
    compacted B1) and one G2 batch of 1 MSM per request (the compacted B2
    query, 2n/3 points); seeded uniform scalars and bases k_i*G."""
    import liblightzk_amd as z
    mix = [12] * 13 + [16] * 10 + [18] * 6 + [20] * 3
    nmax = 3 << 20                       # the merged C MSM of a 2^20 request: 8/3 * 2^20 points
    s = z.gen_field(seed, 0, 2 * nmax, "fr")
    k = z.gen_field(seed, 1, nmax, "fr")
    ds, dk = ctx.upload(s), ctx.upload(k)
    b1, b2 = ctx.alloc(64 * nmax), ctx.alloc(128 * (1 << 20))
    ctx.fixbase_g1(dk, nmax, b1)
    ctx.fixbase_g2(dk, 1 << 20, b2)
    out1, out2 = ctx.alloc(96 * 2 * len(mix)), ctx.alloc(192 * len(mix))
    rng = np.random.default_rng([seed, 9])
    j1, j2 = [], []
    for i, lg in enumerate(mix):
        n = 1 << lg
        for t, m in enumerate((2 * n // 3, 8 * n // 3)):
            j1.append((ds.ptr + 32 * int(rng.integers(0, nmax)), b1.ptr, m, out1.ptr + 96 * (2 * i + t)))
        j2.append((ds.ptr + 32 * int(rng.integers(0, nmax)), b2.ptr, 2 * n // 3, out2.ptr + 192 * i))
    res = {}
    for grp, jobs in ((1, j1), (2, j2)):
        ctx.msm_batch(grp, jobs)   # warm-up (scratch, plans)
        ts = []
        for _ in range(reps):
            ctx.sync()
            t = time.perf_counter()
            ctx.msm_batch(grp, jobs)
            ctx.sync()
            ts.append(time.perf_counter() - t)
        pts = sum(j[2] for j in jobs)
        res[f"g{grp}"] = {"msms": len(jobs), "points": pts, "ms": min(ts) * 1e3, "Mpoints_per_s": pts / min(ts) / 1e6}
    for buf in (ds, dk, b1, b2, out1, out2):
        buf.free()
    return res


def batched_forms(ctx, count=128, seed=4, reps=3):
    """128 G1 MSMs with n = floor(2^(10 + 10U)) and 128 Fr NTTs of 2^k, k
    uniform in 10..20 (seeded stream 6 of seed 4): each set as one batched,
    hipGraph-replayed call against the same instances launched one at a
    time; outputs compared byte for byte between the two."""
    import liblightzk_amd as z
    rng = np.random.default_rng([seed, 6])
    sizes = [int(2 ** (10 + 10 * u)) for u in rng.random(count)]
    nmax = 1 << 20
    s = z.gen_field(seed, 0, 2 * nmax, "fr")
    k = z.gen_field(seed, 1, nmax, "fr")
    d_s, d_k, d_b = ctx.upload(s), ctx.upload(k), ctx.alloc(64 * nmax)
    ctx.fixbase_g1(d_k, nmax, d_b)
    offs = [int(o) for o in rng.integers(0, nmax, size=count)]
    outs_b = ctx.alloc(96 * count)
    outs_1 = ctx.alloc(96 * count)
    jobs = [(d_s.ptr + 32 * o, d_b.ptr, m, outs_b.ptr + 96 * i) for i, (o, m) in enumerate(zip(offs, sizes))]
    ctx.msm_batch(1, jobs, graph=True)   # capture
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        ctx.msm_batch(1, jobs, graph=True)
    ctx.sync()
    tb = (time.perf_counter() - t) / reps
    for i, (o, m) in enumerate(zip(offs, sizes)):   # warm plans / scratch
        ctx.msm_g1(d_s.ptr + 32 * o, d_b, m, outs_1.offset(96 * i))
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        for i, (o, m) in enumerate(zip(offs, sizes)):
            ctx.msm_g1(d_s.ptr + 32 * o, d_b, m, outs_1.offset(96 * i))
    ctx.sync()
    t1 = (time.perf_counter() - t) / reps
    msm_same = bool(np.array_equal(outs_b.download(np.uint64, 12 * count), outs_1.download(np.uint64, 12 * count)))
    pts = sum(sizes)
    # NTTs
    logs = [int(v) for v in rng.integers(10, 21, size=count)]
    tot = sum(1 << lg for lg in logs)
    x = z.gen_field(seed, 8, tot, "fr")
    d_a, d_c = ctx.upload(x), ctx.upload(x)
    base, jobs_n = 0, []
    for lg in logs:
        jobs_n.append((lg, base))
        base += 32 << lg
    ctx.ntt_batch([(d_a.ptr + o, 1 << lg, 0) for lg, o in jobs_n], graph=True)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        ctx.ntt_batch([(d_a.ptr + o, 1 << lg, 0) for lg, o in jobs_n], graph=True)
    ctx.sync()
    nb = (time.perf_counter() - t) / reps
    for lg, o in jobs_n:
        ctx.ntt(d_c.ptr + o, 1 << lg)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        for lg, o in jobs_n:
            ctx.ntt(d_c.ptr + o, 1 << lg)
    ctx.sync()
    n1 = (time.perf_counter() - t) / reps
    ntt_same = bool(np.array_equal(d_a.download(np.uint64, 4 * tot), d_c.download(np.uint64, 4 * tot)))
    bfly = sum((1 << lg) // 2 * lg for lg in logs)
    # one 2^18 transform at a time (the size of a proof's transforms)
    n18 = 1 << 18
    d_18 = ctx.upload(x[: 4 * n18])
    ctx.ntt(d_18, n18)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(20):
        ctx.ntt(d_18, n18)
    ctx.sync()
    t18 = (time.perf_counter() - t) / 20
    d_18.free()
    for b in (d_s, d_k, d_b, outs_b, outs_1, d_a, d_c):
        b.free()
    return {"msm": {"instances": count, "points": pts, "sizes": "floor(2^(10+10U)), seeded",
                    "batched_graph_ms": tb * 1e3, "one_at_a_time_ms": t1 * 1e3,
                    "batched_Mpoints_per_s": pts / tb / 1e6, "one_at_a_time_Mpoints_per_s": pts / t1 / 1e6,
                    "speedup": t1 / tb, "outputs_identical": msm_same},
            "ntt": {"transforms": count, "elements": tot, "sizes": "2^k, k uniform in 10..20, seeded",
                    "batched_graph_ms": nb * 1e3, "one_at_a_time_ms": n1 * 1e3,
                    "batched_Gbutterflies_per_s": bfly / nb / 1e9, "one_at_a_time_Gbutterflies_per_s": bfly / n1 / 1e9,
                    "speedup": n1 / nb, "outputs_identical": ntt_same,
                    "single_2^18_us": t18 * 1e6, "single_2^18_Gbutterflies_per_s": (n18 // 2) * 18 / t18 / 1e9,
                    "roofline": intmul_roofline(bfly * INTMUL_PER_FQ, nb, "batched NTT (whole batch, wall clock)",
                                                {"hbm_achieved_GBps": tot * 64 / nb / 1e9, "hbm_peak_GBps": HBM_PEAK / 1e9,
                                                 "hbm_frac": tot * 64 / nb / HBM_PEAK,
                                                 "algorithmic_bytes": tot * 64,
                                                 "single_2^18_frac": w_ntt(n18) / t18 / INTMUL_PEAK})}}


def large_msm_variants(ctx, reps=2):
    """G2 MSM at n = 2^22 (bases k_i * G2) and the G1 2^24 MSM on sparse scalars
    (70% in {0, 1}), the other large single-MSM shapes of the workload set."""
    import liblightzk_amd as z
    out = {}
    n = 1 << 22
    s, k = z.gen_field(3, 0, n, "fr"), z.gen_field(3, 1, n, "fr")
    d_s, d_k, d_b, d_o = ctx.upload(s), ctx.upload(k), ctx.alloc(128 * n), ctx.alloc(192)
    ctx.fixbase_g2(d_k, n, d_b)
    ctx.msm_g2(d_s, d_b, n, d_o)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        ctx.msm_g2(d_s, d_b, n, d_o)
    ctx.sync()
    el = (time.perf_counter() - t) / reps
    out["g2_msm_2^22_uniform"] = {"ms": el * 1e3, "Mpoints_per_s": n / el / 1e6}
    for b in (d_s, d_k, d_b, d_o):
        b.free()
    n = 1 << 24
    s, k = z.gen_field(3, 0, n, "fr_sparse"), z.gen_field(3, 1, n, "fr")
    d_s, d_k, d_b, d_o = ctx.upload(s), ctx.upload(k), ctx.alloc(64 * n), ctx.alloc(96)
    ctx.fixbase_g1(d_k, n, d_b)
    ctx.msm_g1(d_s, d_b, n, d_o)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(reps):
        ctx.msm_g1(d_s, d_b, n, d_o)
    ctx.sync()
    el = (time.perf_counter() - t) / reps
    st = ctx.msm_last_stats()
    out["g1_msm_2^24_sparse"] = {"ms": el * 1e3, "Mpoints_per_s": n / el / 1e6, "nonzero_digits": st["entries"],
                                 "largest_bucket": st["max_bucket"]}
    for b in (d_s, d_k, d_b, d_o):
        b.free()
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(len(xs) * q))]


def groth16_served(logs, weights, requests, rate_mult, max_batch, seed=5, pipelines=3, levels=1,
                   modes=("sequential", "continuous"), under_load_requests=150):
    """Mixed-size request stream served two ways by one hzk_server holding all
    circuits' keys: continuous batching (mode 0) and the plain sequential
    prover through the same queue (mode 1). Requests arrive as a seeded Poisson
    stream at rate_mult x the plain prover's measured capacity on this mix;
    each carries its own seeded witness (generated before the clock starts).
    Both runs must return the same proof bytes request by request. Then the
    first under_load_requests requests arrive at half the plain capacity
    (the under-load latency point), in both modes, with the same proofs."""
    import threading
    import liblightzk_amd as z
    srv = z.ProofServer(0, max_batch=max_batch, mode="sequential", pipelines=pipelines)
    t = time.perf_counter()
    cids = [srv.add_circuit(40 + lg, lg, 8) for lg in logs]
    if levels > 1:
        for cid in cids:
            srv.set_levels(cid, levels)
    setup_s = time.perf_counter() - t
    log(f"served: setup of {len(logs)} circuits {setup_s:.1f} s")
    rng = np.random.default_rng(seed)
    p = np.array(weights, dtype=float) / sum(weights)
    kinds = rng.choice(len(logs), size=requests, p=p)
    gaps = rng.exponential(1.0, size=requests)
    reqs = [(int(k), srv.witness(cids[k], 1000 + i), z.Groth16.rs(1000 + i)) for i, k in enumerate(kinds)]
    srv.start()
    # plain per-circuit proof time (one warm request each, then 2 timed)
    plain_t = []
    for k in range(len(logs)):
        zz, rs = srv.witness(cids[k], 7), z.Groth16.rs(7)
        srv.wait(srv.submit(cids[k], zz, rs))
        ts = []
        for _ in range(2):
            a = time.perf_counter()
            srv.wait(srv.submit(cids[k], zz, rs))
            ts.append(time.perf_counter() - a)
        plain_t.append(min(ts))
    capacity = 1.0 / float(np.dot(p, plain_t))
    rate = rate_mult * capacity
    arrivals = np.cumsum(gaps / rate)

    def run(mode, count=requests, arr=arrivals):
        srv.set_mode(mode)
        ids = [None] * count
        t0 = time.perf_counter()

        def client():
            for i, (k, zz, rs) in enumerate(reqs[:count]):
                d = t0 + arr[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[i] = srv.submit(cids[k], zz, rs)

        th = threading.Thread(target=client)
        th.start()
        out, lat = [], []
        for i in range(count):
            while ids[i] is None:
                time.sleep(0.0005)
            st, proof, l = srv.wait(ids[i], 600)
            if st != 1:
                raise RuntimeError(f"served request {i} failed: {st} {srv.last_error()}")
            out.append(proof)
            lat.append(l)
        el = time.perf_counter() - t0
        th.join()
        return out, lat, el

    res = {}
    proofs = {}
    for mode in modes:
        st0 = srv.stats()
        proofs[mode], lat, el = run(mode)
        st1 = srv.stats()
        res[mode] = {"proofs_per_s": requests / el, "p50_ms": _pct(lat, 0.5) * 1e3, "p99_ms": _pct(lat, 0.99) * 1e3,
                     "wall_s": el, "steps": st1["steps"] - st0["steps"], "max_inflight": st1["max_inflight"]}
        log(f"served: {mode} {res[mode]['proofs_per_s']:.1f} proofs/s p99 {res[mode]['p99_ms']:.0f} ms")
    # under load: the same stream's first requests at half the plain capacity,
    # where latency is the proving time plus any wait for a batch to form
    under = {}
    if len(res) == 2 and under_load_requests > 0:
        n2 = min(requests, under_load_requests)
        arr2 = np.cumsum(gaps[:n2] / (0.5 * capacity))
        for mode in modes:
            pr, lat, el = run(mode, n2, arr2)
            if not all(np.array_equal(a, b) for a, b in zip(pr, proofs[mode][:n2])):
                raise RuntimeError(f"under-load {mode} proofs differ from the saturated run's")
            under[mode] = {"proofs_per_s": n2 / el, "p50_ms": _pct(lat, 0.5) * 1e3, "p99_ms": _pct(lat, 0.99) * 1e3}
            log(f"served under load (0.5x): {mode} p50 {under[mode]['p50_ms']:.1f} ms p99 {under[mode]['p99_ms']:.1f} ms")
    srv.close()
    if len(res) < 2:
        return {"requests": requests, "key_levels": levels, **res}
    same = all(np.array_equal(a, b) for a, b in zip(proofs["sequential"], proofs["continuous"]))
    return {"circuits": [f"m=2^{lg}" for lg in logs], "weights": list(weights), "requests": requests,
            "arrivals": f"Poisson, seed {seed}, rate {rate:.1f}/s = {rate_mult}x plain capacity {capacity:.1f}/s",
            "plain_ms_per_circuit": [x * 1e3 for x in plain_t], "setup_s": setup_s, "max_batch": max_batch,
            "pipelines": pipelines, "key_levels": levels,
            "sequential": res["sequential"], "continuous": res["continuous"],
            "speedup_proofs_per_s": res["continuous"]["proofs_per_s"] / res["sequential"]["proofs_per_s"],
            "under_load": {"requests": min(requests, under_load_requests), "rate": f"0.5x plain capacity = {0.5 * capacity:.1f}/s",
                           **under} if under else None,
            "proofs_identical_across_modes": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--g16-log-m", type=int, default=18)
    ap.add_argument("--g16-proofs", type=int, default=5)
    ap.add_argument("--g16-levels", type=int, default=1,
                    help="fixed-base level tables per proving key (levels per bucket set; 1 = plain key arrays)")
    ap.add_argument("--serve-logs", default="12,16,18,20")
    ap.add_argument("--serve-weights", default="40,30,20,10")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="length of the served request stream (a finite Poisson stream of this many requests)")
    ap.add_argument("--serve-rate-mult", type=float, default=4.0)
    ap.add_argument("--serve-max-batch", type=int, default=32)
    ap.add_argument("--serve-pipelines", type=int, default=3)
    ap.add_argument("--serve-modes", default="sequential,continuous",
                    help="served-leg modes to run (both are needed for the speedup figure)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--served-batch-reps", type=int, default=3,
                    help="timed runs of the served mix's MSM batches launched alone (0 = skip)")
    ap.add_argument("--no-presets", action="store_true", help="skip the reference's MatMul p = 39/63/100 proofs")
    ap.add_argument("--no-large", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs on one GPU only "
              f"(the MSM would shard by points or windows with one 96-byte exchange per GPU; not built)",
              file=sys.stderr)
        sys.exit(2)

    import liblightzk_amd as z
    ctx = z.Context(0)
    n = 1 << args.log_n
    seed = 3
    scalars = z.gen_field(seed, 0, n, "fr")
    ks = z.gen_field(seed, 1, n, "fr")
    d_s = ctx.upload(scalars)
    d_k = ctx.upload(ks)
    d_b = ctx.alloc(64 * n)
    ctx.fixbase_g1(d_k, n, d_b)
    ctx.sync()
    d_k.free()
    d_out = ctx.alloc(96)

    peak_mad = ctx.roofline_probe("mad")
    hbm_copy = ctx.roofline_probe("copy")
    for _ in range(args.warmup):
        ctx.msm_g1(d_s, d_b, n, d_out)
    ctx.sync()
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.msm_g1(d_s, d_b, n, d_out)
    ctx.sync()
    el = time.perf_counter() - t0
    rep = ctx.timing_report()
    ctx.timing(False)
    st = ctx.msm_last_stats()
    ms_step = el / args.steps * 1e3
    value = n / (el / args.steps) / 1e6

    acc_ms, acc_n = rep.get("msm_accumulate", (0.0, 0))
    acc_avg_s = acc_ms / max(acc_n, 1) * 1e-3
    work = w_msm(n)
    entry_work = st["entries"] * FQ_PER_MADD * INTMUL_PER_FQ
    traffic = None   # HBM bytes per accumulate launch, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    prof = os.path.join(ROOT, "profiles", "msm_g1_2e24_hbm_traffic.txt")
    if args.log_n == 24 and os.path.exists(prof):
        for line in open(prof):
            if line.startswith("k_msm_accumulate"):
                traffic = float(line.split()[-1]) * 1e9
    roofline = intmul_roofline(work, acc_avg_s, "msm_accumulate", {
        "traffic": traffic,
        "traffic_source": "profiles/msm_g1_2e24_hbm_traffic.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)",
        "work_formula": "W_MSM(n) = min_c ceil(255/c) (11 n + 16 2^c) * 128",
        "avg_launch_ms": acc_avg_s * 1e3,
        "measured_peak": peak_mad / 1e12,
        "frac_of_measured_peak": (work / acc_avg_s) / peak_mad if peak_mad and acc_avg_s > 0 else None,
        "measured_peak_source": "v_mad_u64_u32 issue rate measured on this GPU (hzk_roofline_probe)",
        "entry_work_per_launch": entry_work,
        "entry_frac": (entry_work / acc_avg_s) / INTMUL_PEAK if acc_avg_s > 0 else None,
        "entry_formula": "this plan's nonzero bucket entries x 11 x 128",
        "whole_msm_frac": (work / (el / args.steps)) / INTMUL_PEAK,
        "algorithmic_bytes_per_step": n * 96,
        "hbm_frac": n * 96 / (el / args.steps) / HBM_PEAK,
        "hbm_copy_measured_GBps": hbm_copy / 1e9,
    })
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded Fr scalars; bases k_i*G1 generated on the GPU)",
        "config": {"workload": f"g1_msm_2^{args.log_n}_uniform", "n": n, "plan": "built-in autotuned table (liblightzk_amd/tune/mi355x.json)",
                   "window_bits": st["c"], "chunk_K": ctx.msm_tune_get(1, args.log_n)[1],
                   "windows": st["windows"], "bucket_entries": st["entries"], "parallelism": "single-gpu"},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in rep.items()},
        "roofline": roofline,
    }
    log(f"msm: {value:.1f} Mpoints/s")
    # the MSM's last step (Horner over the window sums) runs on the host; what
    # the same ~255 dependent doublings would cost on one device lane
    d1, d2 = ctx.roofline_probe("dbl_chain_g1"), ctx.roofline_probe("dbl_chain_g2")
    out["msm_window_combine"] = {
        "where": "host, after one D2H copy of the window sums (the host worker pool over instances and windows)",
        "device_chain_doubling_us": {"g1": 1e6 / d1, "g2": 1e6 / d2},
        "device_horner_255_doublings_ms": {"g1": 255e3 / d1, "g2": 255e3 / d2},
        "source": "hzk_roofline_probe kinds 4/5: dependent doublings of the 29-bit accumulator, one wave",
    }
    if args.g16_proofs > 0:
        out["groth16_plain"] = groth16_plain(ctx, args.g16_log_m, args.g16_proofs, args.g16_levels)
        log(f"groth16 plain: {out['groth16_plain']['proofs_per_s']:.1f} proofs/s")
    if not args.no_presets:
        out["groth16_reference_presets"] = groth16_reference_presets(ctx)
        log("reference presets: " + ", ".join(f"{k} {v['proofs_per_s']:.1f}/s" for k, v in out["groth16_reference_presets"].items()))
    if args.served_batch_reps > 0:
        out["served_msm_batches"] = served_msm_batches(ctx, args.served_batch_reps)
        log("served MSM batches: " + ", ".join(f"{k} {v['ms']:.1f} ms" for k, v in out["served_msm_batches"].items()))
    if not args.no_batched:
        out["batched"] = batched_forms(ctx)
        log(f"batched: msm x{out['batched']['msm']['speedup']:.2f}, ntt x{out['batched']['ntt']['speedup']:.2f}")
    if not args.no_large:
        out["large_msm"] = large_msm_variants(ctx)
        log(f"large: {json.dumps(out['large_msm'])}")
    if args.serve_requests > 0:
        out["groth16_served"] = groth16_served([int(v) for v in args.serve_logs.split(",")],
                                               [float(v) for v in args.serve_weights.split(",")],
                                               args.serve_requests, args.serve_rate_mult, args.serve_max_batch,
                                               pipelines=args.serve_pipelines, levels=args.g16_levels,
                                               modes=args.serve_modes.split(","))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ctx, scalars, d_b, args.cpu_seconds)
        out["cpu_baseline"]["groth16"] = cpu_prover_baseline(ctx)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
