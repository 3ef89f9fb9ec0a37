"""hzk_msm_g1_noalloc: the G1 MSM on the static device scratch, no allocation.

Two calls in a row through the C ABI, the second on a smaller input, each
equal to the allocating entry point hzk_msm_g1 on the same inputs and to the
oracle's closed form (sum s_i k_i) G for bases P_i = k_i G, bit for bit;
an MSM too large for the static scratch is refused with HZK_ERR_UNSUPPORTED
and leaves the entry point usable."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(ctx, orc, seed, n):
    s, k = orc.gen_fr(seed, 0, n), orc.gen_fr(seed, 1, n)
    ds, dk, db = ctx.upload(s), ctx.upload(k), ctx.alloc(64 * n)
    ctx.fixbase_g1(dk, n, db)
    return s, k, ds, db


def test_noalloc_matches_allocating_entry_point(orc):
    import liblightzk_amd as z
    ctx = z.Context(0)
    try:
        for seed, n in ((71, 1 << 14), (72, 1 << 11)):   # the second call smaller than the first
            s, k, ds, db = _inputs(ctx, orc, seed, n)
            d0, d1 = ctx.alloc(96), ctx.alloc(96)
            ctx.msm_g1_noalloc(ds, db, n, d0)
            ctx.msm_g1(ds, db, n, d1)
            got, want = d0.download(np.uint64, 12), d1.download(np.uint64, 12)
            assert np.array_equal(got, want), n
            assert np.array_equal(got, orc.msm_closed_form(s, k, 1)), n
        # too large for the static scratch: refused, and the entry point still works afterwards
        n = 1 << 20
        s, k, ds, db = _inputs(ctx, orc, 73, n)
        d0 = ctx.alloc(96)
        with pytest.raises(z.HzkError) as e:
            ctx.msm_g1_noalloc(ds, db, n, d0)
        assert "rc=-" in str(e.value) and "scratch" in str(e.value)
        s, k, ds, db = _inputs(ctx, orc, 74, 300)
        ctx.msm_g1_noalloc(ds, db, 300, d0)
        assert np.array_equal(d0.download(np.uint64, 12), orc.msm_closed_form(s, k, 1))
    finally:
        ctx.close()
