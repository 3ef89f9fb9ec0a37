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


BATCH_PLANS = [(6, 16, 4), (9, 64, 32), (11, 32, 64), (13, 128, 16)]


@pytest.mark.parametrize("plan", BATCH_PLANS)
def test_batch_regime_plans_exact(ctx, orc, data, plan):
    """Wide-batch plans (forced regime) give the closed form, in a batch and
    alone, for G1 and G2."""
    n, s, sp, k, b1, b2 = data
    c, K, g = plan
    lg = n.bit_length() - 1
    ctx.set_msm_regime(1)
    try:
        for group, bases, width in ((1, b1, 12), (2, b2, 24)):
            ctx.msm_batch_tune_set(group, lg, c, K, g)
            assert ctx.msm_batch_tune_get(group, lg) == (c, K, g)
            db = ctx.upload(bases)
            for sc in (s, sp):
                ds = ctx.upload(sc)
                outs = [ctx.alloc(8 * width) for _ in range(3)]
                ctx.msm_batch(group, [(ds, db, n, outs[0]), (ds, db, 1000, outs[1]), (ds, db, n - 1, outs[2])])
                assert np.array_equal(outs[0].download(np.uint64, width), orc.msm_closed_form(sc, k, group)), plan
                assert np.array_equal(outs[1].download(np.uint64, width),
                                      orc.msm_closed_form(sc[: 4 * 1000].copy(), k[: 4 * 1000].copy(), group))
                assert np.array_equal(outs[2].download(np.uint64, width),
                                      orc.msm_closed_form(sc[: 4 * (n - 1)].copy(), k[: 4 * (n - 1)].copy(), group))
                d = ctx.alloc(8 * width)
                (ctx.msm_g1 if group == 1 else ctx.msm_g2)(ds, db, n, d)
                assert np.array_equal(d.download(np.uint64, width), orc.msm_closed_form(sc, k, group)), plan
    finally:
        ctx.set_msm_regime(-1)
        ctx.msm_batch_tune_set(1, lg, 0)
        ctx.msm_batch_tune_set(2, lg, 0)


def test_wide_batch_takes_batch_plans_same_results(orc):
    """Two 2^21-point G1 MSMs (2^26 digits under the single plans) run as a
    batch take the class's wide-batch plan, and give the same points as each
    MSM run alone under the single-MSM plan."""
    import liblightzk_amd as z
    n = 1 << 21
    with z.Context(0) as c2:
        k = z.gen_field(9, 1, n)
        s = z.gen_field(9, 0, n + 64)
        dk, ds = c2.upload(k), c2.upload(s)
        db = c2.alloc(64 * n)
        c2.fixbase_g1(dk, n, db)
        c2.msm_batch_tune_set(1, 21, 14, 64, 32)
        single_c = c2.msm_tune_get(1, 21)[0]
        assert single_c != 14
        outs = [c2.alloc(96), c2.alloc(96)]
        c2.msm_batch(1, [(ds, db, n, outs[0]), (ds.offset(32 * 64), db, n, outs[1])])
        assert c2.msm_last_stats()["c"] == 14
        batch = [o.download(np.uint64, 12) for o in outs]
        for i, off in enumerate((0, 32 * 64)):
            d = c2.alloc(96)
            c2.msm_g1(ds.offset(off), db, n, d)
            assert c2.msm_last_stats()["c"] == single_c
            assert np.array_equal(d.download(np.uint64, 12), batch[i])
        # and one of them against the closed form
        assert np.array_equal(batch[0], orc.msm_closed_form(s[: 4 * n].copy(), k, 1))


def test_batch_autotune_small_classes(orc, data, tmp_path):
    import liblightzk_amd as z
    from liblightzk_amd import tune
    n, s, sp, k, b1, b2 = data
    with z.Context(0) as c2:
        res = c2.msm_batch_autotune(1, 8, 10, budget_ms=3000)
        for lg, (c, K, g) in res.items():
            assert 4 <= c <= 16 and K % 4 == 0 and g & (g - 1) == 0 and g <= 1 << (c - 1), (lg, c, K, g)
        c2.set_msm_regime(1)
        ds, db = c2.upload(s), c2.upload(b1)
        outs = [c2.alloc(96) for _ in range(3)]
        c2.msm_batch(1, [(ds, db, 300, outs[0]), (ds, db, 700, outs[1]), (ds, db, 1500, outs[2])])
        for m, o in zip((300, 700, 1500), outs):
            assert np.array_equal(o.download(np.uint64, 12), orc.msm_closed_form(s[: 4 * m].copy(), k[: 4 * m].copy(), 1))
        path = str(tmp_path / "plans.json")
        tune.save({"device": "gfx950", "g1_batch": {str(lg): list(v) for lg, v in res.items()}}, path)
    with z.Context(0) as c3:
        assert tune.load(c3, path) == len(res)
        for lg, v in res.items():
            assert c3.msm_batch_tune_get(1, lg) == tuple(v)


@pytest.mark.parametrize("c,segment", [(7, 128), (5, 4096), (4, 16)])
def test_forced_segment_longer_than_window(ctx, orc, data, c, segment):
    """hzk_set_tuning with a run length above a window's 2^(c-1) buckets is
    clamped to the window: the MSM stays exact (it returned infinity before)."""
    n, s, _, k, b1, b2 = data
    ctx.set_tuning(c, segment)
    try:
        ds, db1, db2, d1, d2 = ctx.upload(s), ctx.upload(b1), ctx.upload(b2), ctx.alloc(96), ctx.alloc(192)
        ctx.msm_g1(ds, db1, n, d1)
        ctx.msm_g2(ds, db2, n, d2)
        assert np.array_equal(d1.download(np.uint64, 12), orc.msm_closed_form(s, k, 1))
        assert np.array_equal(d2.download(np.uint64, 24), orc.msm_closed_form(s, k, 2))
    finally:
        ctx.set_tuning(0, 0)
