"""GPU Pippenger MSM (hzk_msm_g1 / hzk_msm_g2) vs the CPU oracle, bit-exact.

Semantics of the reference operator (include/operator/msm.cuh:210-248, checked
there by test/msm-check.cu:43-57 against libff): dest = sum s_i * P_i, written
as a normalised Jacobian point. Pinned by the oracle's Pippenger, its naive
double-and-add sum and the closed form (sum s_i k_i) G for P_i = k_i G; edge
cases: n = 1, all-zero and all-one scalars, repeated bases (doubling inside a
bucket), infinity bases, P + (-P), the 70%-{0,1} sparse witness (one hot
bucket, multi-level segment merge) and every supported window size.
"""
import numpy as np
import pytest

import pymodel as pm

pytestmark = pytest.mark.gpu


def _case(orc, seed, n, sparse=False):
    s = orc.gen_fr(seed, 0, n, sparse=sparse)
    k = orc.gen_fr(seed, 1, n)
    return s, k


def _run_g1(ctx, s, bases, win=0):
    n = s.size // 4
    ds, db, dd = ctx.upload(s), ctx.upload(bases), ctx.alloc(96)
    ctx.msm_g1(ds, db, n, dd, win)
    return dd.download(np.uint64, 12)


def _run_g2(ctx, s, bases, win=0):
    n = s.size // 4
    ds, db, dd = ctx.upload(s), ctx.upload(bases), ctx.alloc(192)
    ctx.msm_g2(ds, db, n, dd, win)
    return dd.download(np.uint64, 24)


@pytest.fixture(scope="module")
def g1_data(orc):
    n = 1 << 14
    s, k = _case(orc, 1, n)
    return s, k, orc.g1_mul_gen(k)


@pytest.mark.parametrize("n", [1, 2, 3, 64, 1000, 1 << 14])
def test_g1_random_vs_oracle(ctx, orc, g1_data, n):
    s, k, bases = g1_data
    s, k, bases = s[: 4 * n].copy(), k[: 4 * n].copy(), bases[: 8 * n].copy()
    got = _run_g1(ctx, s, bases)
    assert np.array_equal(got, orc.msm_g1(s, bases, 1)), n
    assert np.array_equal(got, orc.msm_closed_form(s, k, 1)), n


@pytest.mark.parametrize("win", [4, 5, 8, 11, 12, 13, 15, 16])
def test_g1_every_window(ctx, orc, g1_data, win):
    s, k, bases = g1_data
    n = 5000
    s, k, bases = s[: 4 * n].copy(), k[: 4 * n].copy(), bases[: 8 * n].copy()
    assert np.array_equal(_run_g1(ctx, s, bases, win), orc.msm_closed_form(s, k, 1)), win


def test_g1_special_scalars_and_bases(ctx, orc, g1_data):
    s, k, bases = g1_data
    n = 3000
    bases = bases[: 8 * n].copy()
    zero = np.zeros(4 * n, dtype=np.uint64)
    got = _run_g1(ctx, zero, bases)
    assert pm.g1_jac_from_buf(got) is None
    assert np.array_equal(got, orc.msm_g1(zero, bases, 1))   # libff infinity (0, 1, 0)
    one = np.tile(pm.int_to_limbs(pm.MONT % pm.R), n)
    assert np.array_equal(_run_g1(ctx, one, bases), orc.msm_g1(one, bases, 1))
    # repeated bases, a few infinity bases, P and -P in the same MSM
    rep = np.tile(bases[: 8 * 7], n // 7 + 1)[: 8 * n].copy()
    rep[8 * 3: 8 * 4] = 0
    neg = pm.g1_to_aff_buf(pm.G1C.neg(pm.g1_aff_from_buf(rep[:8])))
    rep[8 * 10: 8 * 11] = neg
    sc = s[: 4 * n].copy()
    sc[4 * 10: 4 * 11] = sc[0:4]   # s*P + s*(-P) cancels
    for win in (0, 4, 9):
        assert np.array_equal(_run_g1(ctx, sc, rep, win), orc.msm_g1(sc, rep, 1)), win


def test_g1_sparse_hot_bucket(ctx, orc):
    n = 1 << 15
    s, k = _case(orc, 3, n, sparse=True)
    bases = orc.g1_mul_gen(k)
    want = orc.msm_closed_form(s, k, 1)
    for win in (0, 4, 8):   # c = 4: ~2^15/8 entries per bucket -> several merge levels
        assert np.array_equal(_run_g1(ctx, s, bases, win), want), win


def test_g1_deterministic(ctx, orc, g1_data):
    s, k, bases = g1_data
    a = _run_g1(ctx, s, bases)
    b = _run_g1(ctx, s, bases)
    assert np.array_equal(a, b)


def test_g1_naive_semantics_small(ctx, orc):
    """test/msm-check.cu:46-55: per-thread s_i * P_i accumulators (oracle method 0)."""
    s, k = _case(orc, 9, 300)
    bases = orc.g1_mul_gen(k)
    assert np.array_equal(_run_g1(ctx, s, bases), orc.msm_g1(s, bases, 0))


@pytest.mark.parametrize("n", [1, 50, 777])
def test_g2_vs_oracle(ctx, orc, n):
    s, k = _case(orc, 4, n)
    bases = orc.g2_mul_gen(k)
    got = _run_g2(ctx, s, bases)
    assert np.array_equal(got, orc.msm_g2(s, bases, 1)), n
    assert np.array_equal(got, orc.msm_closed_form(s, k, 2)), n


def test_g2_special(ctx, orc):
    n = 400
    s, k = _case(orc, 5, n, sparse=True)
    bases = orc.g2_mul_gen(k)
    bases[16 * 7: 16 * 8] = 0
    rep = bases.copy()
    rep[16 * 20: 16 * 40] = np.tile(bases[16 * 19: 16 * 20], 20)
    for win in (0, 4, 7):
        assert np.array_equal(_run_g2(ctx, s, rep, win), orc.msm_g2(s, rep, 1)), win


def test_msm_rejects_bad_arguments(ctx):
    from liblightzk_amd import HzkError
    d = ctx.alloc(4096)
    with pytest.raises(HzkError):
        ctx.msm_g1(d, d, 0, d)
    with pytest.raises(HzkError):
        ctx.msm_g1(d, d, 4, d, win_siz=17)


@pytest.mark.parametrize("n", [1, 33, 4097])
def test_fixbase_g1_g2(ctx, orc, n):
    """hzk_fixbase_g1/_g2: k_i * G normalised to affine, bit-exact with the oracle."""
    k = orc.gen_fr(12, 1, n)
    k[:4] = 0                                          # 0 * G = infinity (zero bytes)
    if n > 1:
        k[4:8] = pm.int_to_limbs(pm.MONT % pm.R)       # 1 * G
    dk = ctx.upload(k)
    o1, o2 = ctx.alloc(64 * n), ctx.alloc(128 * n)
    ctx.fixbase_g1(dk, n, o1)
    ctx.fixbase_g2(dk, n, o2)
    assert np.array_equal(o1.download(np.uint64, 8 * n), orc.g1_mul_gen(k))
    assert np.array_equal(o2.download(np.uint64, 16 * n), orc.g2_mul_gen(k))
    assert np.array_equal(dk.download(np.uint64, 4 * n), k)   # scalars untouched


def _batch_case(orc, seed, sizes, group):
    tot = sum(sizes)
    s = orc.gen_fr(seed, 0, tot)
    k = orc.gen_fr(seed, 1, tot)
    bases = orc.g1_mul_gen(k) if group == 1 else orc.g2_mul_gen(k)
    return s, k, bases


@pytest.mark.parametrize("graph", [False, True])
def test_g1_batch_uneven_sizes(ctx, orc, graph):
    """Batched MSMs (one launch per stage over all instances, optionally one
    hipGraph): every result equals the same instance run alone and the closed form."""
    rng = np.random.default_rng(7)
    sizes = [int(2 ** rng.uniform(4, 14)) for _ in range(24)] + [1, 2]
    s, k, bases = _batch_case(orc, 21, sizes, 1)
    ds, db = ctx.upload(s), ctx.upload(bases)
    outs = ctx.alloc(96 * len(sizes))
    for rep in range(2):   # second pass replays the captured graph with new scalars
        if rep == 1:
            s = orc.gen_fr(22, 0, sum(sizes))
            ds.upload(s)
        jobs, off = [], 0
        for i, n in enumerate(sizes):
            jobs.append((ds.offset(32 * off), db.offset(64 * off), n, outs.offset(96 * i)))
            off += n
        ctx.msm_batch(1, jobs, graph=graph)
        got = outs.download(np.uint64, 12 * len(sizes))
        off = 0
        for i, n in enumerate(sizes):
            want = orc.msm_closed_form(s[4 * off: 4 * (off + n)].copy(), k[4 * off: 4 * (off + n)].copy(), 1)
            assert np.array_equal(got[12 * i: 12 * i + 12], want), (rep, i, n)
            single = _run_g1(ctx, s[4 * off: 4 * (off + n)].copy(), bases[8 * off: 8 * (off + n)].copy())
            assert np.array_equal(single, want), (rep, i, n)
            off += n


def test_g2_batch_host_dest(ctx, orc):
    sizes = [5, 300, 1000, 64]
    s, k, bases = _batch_case(orc, 23, sizes, 2)
    ds, db = ctx.upload(s), ctx.upload(bases)
    host_out = [np.zeros(24, dtype=np.uint64) for _ in sizes]
    jobs, off = [], 0
    for i, n in enumerate(sizes):
        jobs.append((ds.offset(32 * off), db.offset(128 * off), n, host_out[i].ctypes.data))
        off += n
    ctx.msm_batch(2, jobs, host_dest=True)
    off = 0
    for i, n in enumerate(sizes):
        want = orc.msm_closed_form(s[4 * off: 4 * (off + n)].copy(), k[4 * off: 4 * (off + n)].copy(), 2)
        assert np.array_equal(host_out[i], want), (i, n)
        off += n


@pytest.mark.parametrize("graph", [False, True])
def test_device_combine_matches_host_combine(ctx, orc, graph):
    """HZK_BATCH_DEVICE_COMBINE (the window combine as a kernel, no host
    synchronisation inside the call): the same bytes as the host combine and
    the closed form, G1 and G2, including an all-zero instance (the point at
    infinity) and a replayed graph."""
    rng = np.random.default_rng(11)
    for group in (1, 2):
        sizes = [int(2 ** rng.uniform(3, 12)) for _ in range(10)] + [1, 37]
        s, k, bases = _batch_case(orc, 31 + group, sizes, group)
        s[: 4 * sizes[0]] = 0
        pw, bw = (96, 64) if group == 1 else (192, 128)
        ds, db = ctx.upload(s), ctx.upload(bases)
        out_h, out_d = ctx.alloc(pw * len(sizes)), ctx.alloc(pw * len(sizes))

        def jobs(out):
            js, off = [], 0
            for i, n in enumerate(sizes):
                js.append((ds.offset(32 * off), db.offset(bw * off), n, out.offset(pw * i)))
                off += n
            return js

        ctx.msm_batch(group, jobs(out_h), graph=graph)
        want = out_h.download(np.uint64, pw // 8 * len(sizes))
        for rep in range(2):
            out_d.zero()
            ctx.msm_batch(group, jobs(out_d), graph=graph, device_combine=True)
            ctx.sync()
            got = out_d.download(np.uint64, pw // 8 * len(sizes))
            assert np.array_equal(got, want), (group, rep)
        off = 0
        for i, n in enumerate(sizes):
            cf = orc.msm_closed_form(s[4 * off: 4 * (off + n)].copy(), k[4 * off: 4 * (off + n)].copy(), group)
            assert np.array_equal(want[pw // 8 * i: pw // 8 * (i + 1)], cf), (group, i, n)
            off += n
        for b in (ds, db, out_h, out_d):
            b.free()


def test_device_combine_rejects_host_dest(ctx):
    import liblightzk_amd as z
    d = ctx.alloc(64)
    with pytest.raises(z.HzkError):
        ctx.msm_batch(1, [(d.ptr, d.ptr, 1, 0)], host_dest=True, device_combine=True)
    d.free()
