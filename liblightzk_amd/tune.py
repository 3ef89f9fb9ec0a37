"""Persisted MSM plans (window bits c, chunk length K, run length g) per group
and size class, produced by the library's autotuner (hzk_msm_autotune) on a
given GPU and loaded into a context before use.

File format (JSON): {"device": str, "g1": {"<log_n>": [c, K, g], ...},
"g2": {...}, "g1_batch": {...}, "g2_batch": {...}}; the *_batch tables hold
the plans of the wide-batch (throughput) regime."""
import json
import os

DEFAULT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune", "mi355x.json")


def autotune(ctx, classes_g1=(10, 24), classes_g2=(10, 22), budget_ms=4000.0, batch=False):
    """Tune the single-MSM regime (batch=False) or the wide-batch regime."""
    out = {"device": "gfx950"}
    for group, (lo, hi), key in ((1, classes_g1, "g1"), (2, classes_g2, "g2")):
        if lo is None:
            continue
        if batch:
            res = ctx.msm_batch_autotune(group, lo, hi, budget_ms)
            key += "_batch"
        else:
            res = ctx.msm_autotune(group, lo, hi, budget_ms)
        out[key] = {str(lg): list(v) for lg, v in res.items()}
    return out


def save(table, path=DEFAULT):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(ctx, path=DEFAULT):
    """Load a plan table into ctx; returns the number of entries (0 if no file)."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        table = json.load(f)
    n = 0
    for group, key in ((1, "g1"), (2, "g2")):
        for lg, (c, K, g) in table.get(key, {}).items():
            ctx.msm_tune_set(group, int(lg), int(c), int(K), int(g))
            n += 1
        for lg, (c, K, g) in table.get(key + "_batch", {}).items():
            ctx.msm_batch_tune_set(group, int(lg), int(c), int(K), int(g))
            n += 1
    return n
