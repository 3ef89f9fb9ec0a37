"""MSM plans and the autotuner.

Results must not depend on the plan: every (window bits c, chunk length K,
run length g) in the tuner's search space gives the oracle's closed form
(sum s_i k_i) G bit for bit, for G1 and G2, uniform and sparse (hot bucket)
scalars, alone and in a batch. The autotuner stores a valid plan per class
and MSMs after it stay exact; a plan table round-trips through JSON."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PLANS = [(4, 4, 1), (5, 8, 2), (7, 4, 8), (9, 64, 4), (12, 16, 32), (13, 256, 1), (16, 4096, 16), (16, 4, 2)]


@pytest.fixture(scope="module")
def data(orc):
    n = 3000
    s = orc.gen_fr(61, 0, n)
    sp = orc.gen_fr(61, 0, n, sparse=True)
    k = orc.gen_fr(61, 1, n)
    return n, s, sp, k, orc.g1_mul_gen(k), orc.g2_mul_gen(k)


@pytest.mark.parametrize("plan", PLANS)
def test_plan_space_g1_g2(ctx, orc, data, plan):
    n, s, sp, k, b1, b2 = data
    c, K, g = plan
    lg = n.bit_length() - 1
    ctx.msm_tune_set(1, lg, c, K, g)
    ctx.msm_tune_set(2, lg, c, K, g)
    try:
        assert ctx.msm_tune_get(1, lg) == (c, K, g)
        db1, db2 = ctx.upload(b1), ctx.upload(b2)
        for sc in (s, sp):
            ds, d1, d2 = ctx.upload(sc), ctx.alloc(96), ctx.alloc(192)
            ctx.msm_g1(ds, db1, n, d1)
            ctx.msm_g2(ds, db2, n, d2)
            assert np.array_equal(d1.download(np.uint64, 12), orc.msm_closed_form(sc, k, 1)), plan
            assert np.array_equal(d2.download(np.uint64, 24), orc.msm_closed_form(sc, k, 2)), plan
    finally:
        ctx.msm_tune_set(1, lg, 0)
        ctx.msm_tune_set(2, lg, 0)


def test_tune_set_rejects_bad_plans(ctx):
    from liblightzk_amd import HzkError
    for bad in ((1, 10, 3, 8, 4), (1, 10, 17, 8, 4), (1, 10, 8, 6, 4), (1, 10, 8, 8, 3), (3, 10, 8, 8, 4),
                (1, 10, 4, 8, 16)):
        with pytest.raises(HzkError):
            ctx.msm_tune_set(*bad)


def test_autotune_then_exact(orc, data, tmp_path):
    import liblightzk_amd as z
    from liblightzk_amd import tune
    n, s, sp, k, b1, b2 = data
    with z.Context(0) as c2:
        res = c2.msm_autotune(1, 8, 12, budget_ms=3000)
        for lg, (c, K, g) in res.items():
            assert 4 <= c <= 16 and K % 4 == 0 and g & (g - 1) == 0 and g <= 1 << (c - 1), (lg, c, K, g)
        ds, db, dd = c2.upload(s), c2.upload(b1), c2.alloc(96)
        c2.msm_g1(ds, db, n, dd)
        assert np.array_equal(dd.download(np.uint64, 12), orc.msm_closed_form(s, k, 1))
        # batch of several classes with tuned plans
        sizes = [300, 1000, 2000, 3000]
        descs, outs = [], []
        for m in sizes:
            o = c2.alloc(96)
            outs.append(o)
            descs.append((ds, db, m, o))
        c2.msm_batch(1, descs)
        for m, o in zip(sizes, outs):
            assert np.array_equal(o.download(np.uint64, 12), orc.msm_closed_form(s[: 4 * m].copy(), k[: 4 * m].copy(), 1))
        path = str(tmp_path / "plans.json")
        tune.save({"device": "gfx950", "g1": {str(lg): list(v) for lg, v in res.items()}, "g2": {}}, path)
    with z.Context(0) as c3:
        assert tune.load(c3, path) == len(res)
        for lg, v in res.items():
            assert c3.msm_tune_get(1, lg) == tuple(v)


@pytest.mark.parametrize("c", [6, 11, 13, 16])
def test_level_tables_single_bucket_set(ctx, orc, data, c):
    """MSM over a fixed-base level table (all windows in one bucket set) equals
    the closed form for G1 and G2, uniform and sparse scalars, alone and in a
    batch mixed with plain instances, with and without hipGraph."""
    n, s, sp, k, b1, b2 = data
    L = (255 + c - 1) // c
    for group, bases, width in ((1, b1, 12), (2, b2, 24)):
        db = ctx.upload(bases)
        tab = ctx.alloc(bases.nbytes * L)
        ctx.bases_levels(group, c, db, n, tab)
        layout_levels = 2
        for sc in (s, sp):
            ds, d = ctx.upload(sc), ctx.alloc(8 * width)
            (ctx.msm_g1 if group == 1 else ctx.msm_g2)(ds, tab, n, d, c, layout_levels)
            assert np.array_equal(d.download(np.uint64, width), orc.msm_closed_form(sc, k, group)), (group, c)
        ds = ctx.upload(s)
        outs = [ctx.alloc(8 * width) for _ in range(3)]
        jobs = [(ds, tab, n, outs[0], c, L), (ds, db, n, outs[1]), (ds, db, 1000, outs[2])]
        part = ctx.alloc(bases.nbytes * 3)          # 3-level table: bucket sets of 3 windows
        ctx.bases_levels(group, c, db, n, part, levels=3)
        jobs.append((ds, part, n, ctx.alloc(8 * width), c, 3))
        outs.append(jobs[-1][3])
        for graph in (False, True, True):
            ctx.msm_batch(group, jobs, graph=graph)
            assert np.array_equal(outs[3].download(np.uint64, width), orc.msm_closed_form(s, k, group))
            assert np.array_equal(outs[0].download(np.uint64, width), orc.msm_closed_form(s, k, group))
            assert np.array_equal(outs[1].download(np.uint64, width), orc.msm_closed_form(s, k, group))
            assert np.array_equal(outs[2].download(np.uint64, width),
                                  orc.msm_closed_form(s[: 4 * 1000].copy(), k[: 4 * 1000].copy(), group))
