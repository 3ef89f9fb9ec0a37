#!/usr/bin/env python3
"""Benchmark: G1 MSM throughput at n = 2^24 on one MI355X.

Workload (seeded, synthetic): scalars uniform in Fr (stream 0 of seed 3),
bases P_i = k_i * G1 with k_i uniform in Fr (stream 1), generated on the GPU
by hzk_fixbase_g1 and resident in HBM before timing. One step = one full
Pippenger MSM (hzk_msm_g1: recode, LDS counting sort, bucket accumulation,
merge, bucket reduction, window combine, normalisation) over all 2^24 points.

Prints ONE JSON line. `value` is Mpoints/s (n / seconds per step).
`roofline` is for the dominant kernel (msm_accumulate): algorithmic 32x32-bit
integer multiplies per launch = entries * 11 * 128 (one mixed addition of
11 Fq products per (point, window) entry, 128 32x32 products per Fq product),
divided by the kernel's average launch time from HIP events on the
context's stream, against the measured v_mad_u64_u32 issue rate.
`cpu_baseline` times the oracle's naive MSM (per-thread s_i * P_i
accumulators, test/msm-check.cu:46-55 semantics) on a bounded sample of the
same inputs on this host's cores.
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


def w_msm(n):
    """Algorithmic int-mul count of a G1 MSM: min_c ceil(255/c) (11 n + 16 2^c) M_Fq."""
    best = None
    for c in range(1, 25):
        w = math.ceil(255 / c) * (11 * n + 16 * 2 ** c) * INTMUL_PER_FQ
        best = w if best is None or w < best else best
    return best


def cpu_baseline(ctx, scalars_host, bases_dev, budget_s):
    """Oracle naive MSM on a bounded sample of the same inputs (the only oracle use here)."""
    from oracle import orc
    nthreads = orc.lib().orc_num_threads()
    n0 = 1 << 10
    t = time.perf_counter()
    orc.msm_g1(scalars_host[: 4 * n0].copy(), bases_dev.download(np.uint64, 8 * n0), 0)
    per_pt = (time.perf_counter() - t) / n0
    n_s = int(min(1 << 20, max(n0, budget_s / max(per_pt, 1e-9))))
    s = scalars_host[: 4 * n_s].copy()
    b = bases_dev.download(np.uint64, 8 * n_s)
    t = time.perf_counter()
    orc.msm_g1(s, b, 0)
    el = time.perf_counter() - t
    return {"value": n_s / el / 1e6, "unit": "Mpoints/s", "cores": nthreads, "kind": "port",
            "sample": f"first {n_s} points of the same workload, naive per-thread double-and-add MSM "
                      f"(test/msm-check.cu:46-55 semantics), {el:.1f} s"}


def groth16_plain(ctx, log_m, proofs):
    """Plain one-proof-at-a-time GPU prover on the synthetic circuit (seed 2)."""
    import liblightzk_amd as z
    t = time.perf_counter()
    g = z.Groth16(ctx, seed=2, log_m=log_m, num_public=8)
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
    g.close()
    lat.sort()
    return {"config": f"m=2^{log_m} multiplication-gate circuit, {g.info['num_vars']} variables",
            "proofs": proofs, "proofs_per_s": proofs / el, "p50_ms": lat[len(lat) // 2] * 1e3,
            "p99_ms": lat[min(len(lat) - 1, int(len(lat) * 0.99))] * 1e3, "setup_s": setup_s,
            "kernels_ms_per_proof": {k: v[0] / proofs for k, v in rep.items()}}


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
    work = st["entries"] * FQ_PER_MADD * INTMUL_PER_FQ
    achieved = work / acc_avg_s if acc_avg_s > 0 else 0.0
    roofline = {
        "bound": "valu-intmul",
        "kernel": "msm_accumulate",
        "achieved": achieved / 1e9,
        "peak": peak_mad / 1e9,
        "unit": "G int-mul/s",
        "frac": achieved / peak_mad if peak_mad else None,
        "traffic": None,
        "peak_source": "measured v_mad_u64_u32 issue rate on this GPU (hzk_roofline_probe)",
        "work_per_launch": work,
        "avg_launch_ms": acc_avg_s * 1e3,
        "msm_intmul_frac": (w_msm(n) / (el / args.steps)) / peak_mad if peak_mad else None,
    }
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
        "config": {"workload": f"g1_msm_2^{args.log_n}_uniform", "n": n, "window_bits": st["c"],
                   "windows": st["windows"], "bucket_entries": st["entries"], "parallelism": "single-gpu"},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in rep.items()},
        "roofline": roofline,
    }
    if args.g16_proofs > 0:
        out["groth16_plain"] = groth16_plain(ctx, args.g16_log_m, args.g16_proofs)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ctx, scalars, d_b, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
