"""Oracle parity at the full size of every benchmarked workload, through the HIP path.

The reference checks its MSM operator at full test size against its CPU path
(test/msm-check.cu:43-57, n = 2^20: per-thread double-and-add accumulators,
then `==`). Here each workload runs at the size it is measured at and is
compared with the oracle on the same seeded bytes:

* G1 MSM, n = 2^16 (seed 1): the oracle's CPU Pippenger, its closed form and
  the GPU fixed-base bases against the oracle's k_i G.
* G1 MSM, n = 2^24, uniform and 70%-{0,1} scalars, and G2 MSM, n = 2^22
  (seed 3): the closed form (sum s_i k_i mod r) G for bases P_i = k_i G, one
  scalar multiplication whatever n is, plus sampled bases against the oracle.
* one Groth16 proof at m = 2^18 (seed 2): the oracle's trapdoor closed form,
  byte for byte, and the oracle's pairing verifier.
* 128 G1 MSMs with n = floor(2^(10+10U)) in one hipGraph-replayed batch
  (seed 4): the closed form per instance on every replay; 128 Fr NTTs of 2^k,
  k uniform in 10..20, in one batch: equal to their plain runs, the oracle's
  NTT for every transform up to 2^12 and one 2^18 transform, and the inverse
  batch restores every input.
* a proof server holding circuits of m = 2^12, 2^16, 2^18 and 2^20 (seeds
  40 + log m) serving a 48-request mix weighted 40/30/20/10: every proof equal
  to the trapdoor closed form, one proof per size pairing-verified.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu_bases(ctx, k, n, group):
    dk = ctx.upload(k)
    db = ctx.alloc((64 if group == 1 else 128) * n)
    (ctx.fixbase_g1 if group == 1 else ctx.fixbase_g2)(dk, n, db)
    dk.free()
    return db


def _sampled_bases_match(orc, db, k, n, group, count=256, seed=0):
    """Bases at `count` seeded indices (first and last included) equal the oracle's k_i G."""
    w = 8 if group == 1 else 16
    idx = np.unique(np.concatenate([[0, n - 1], np.random.default_rng(seed).integers(0, n, count)]))
    ks = np.concatenate([k[4 * i: 4 * i + 4] for i in idx])
    want = (orc.g1_mul_gen if group == 1 else orc.g2_mul_gen)(ks)
    got = np.concatenate([db.download(np.uint64, w, offset=8 * w * int(i)) for i in idx])
    return np.array_equal(got, want)


def _msm(ctx, group, ds, db, n):
    dd = ctx.alloc(96 if group == 1 else 192)
    (ctx.msm_g1 if group == 1 else ctx.msm_g2)(ds, db, n, dd)
    out = dd.download(np.uint64, 12 if group == 1 else 24)
    dd.free()
    return out


def test_g1_msm_2e16_seed1(ctx, orc):
    n = 1 << 16
    s, k = orc.gen_fr(1, 0, n), orc.gen_fr(1, 1, n)
    db = _gpu_bases(ctx, k, n, 1)
    bases = db.download(np.uint64, 8 * n)
    assert np.array_equal(bases, orc.g1_mul_gen(k))
    ds = ctx.upload(s)
    got = _msm(ctx, 1, ds, db, n)
    assert np.array_equal(got, orc.msm_g1(s, bases, 1))
    assert np.array_equal(got, orc.msm_closed_form(s, k, 1))


@pytest.mark.parametrize("sparse", [False, True], ids=["uniform", "sparse70"])
def test_g1_msm_2e24_seed3(ctx, orc, sparse):
    import liblightzk_amd as z
    n = 1 << 24
    s, k = orc.gen_fr(3, 0, n, sparse=sparse), orc.gen_fr(3, 1, n)
    assert np.array_equal(s[: 4 << 12], z.gen_field(3, 0, 1 << 12, "fr_sparse" if sparse else "fr"))
    db = _gpu_bases(ctx, k, n, 1)
    assert _sampled_bases_match(orc, db, k, n, 1)
    ds = ctx.upload(s)
    got = _msm(ctx, 1, ds, db, n)
    assert np.array_equal(got, orc.msm_closed_form(s, k, 1))
    st = ctx.msm_last_stats()
    assert st["entries"] > 0
    ds.free()
    db.free()


def test_g2_msm_2e22_seed3(ctx, orc):
    n = 1 << 22
    s, k = orc.gen_fr(3, 0, n), orc.gen_fr(3, 1, n)
    db = _gpu_bases(ctx, k, n, 2)
    assert _sampled_bases_match(orc, db, k, n, 2, count=64)
    ds = ctx.upload(s)
    got = _msm(ctx, 2, ds, db, n)
    assert np.array_equal(got, orc.msm_closed_form(s, k, 2))
    ds.free()
    db.free()


def test_groth16_2e18_seed2(ctx, orc):
    import liblightzk_amd as z
    g = z.Groth16(ctx, seed=2, log_m=18, num_public=8)
    o = orc.Groth16(seed=2, log_m=18, num_public=8)
    assert g.info["M"] == o.sizes["M"] == 1 << 18 and g.info["num_vars"] == o.sizes["num_vars"]
    for wseed in (11, 12):
        zz = g.witness(wseed)
        assert np.array_equal(zz, o.witness(wseed))
        rs = g.rs(wseed)
        p = g.prove(zz, rs)
        assert np.array_equal(p, o.prove_trapdoor(zz, rs)), wseed
        assert o.verify(zz[: 4 * 9], p), wseed
    bad = p.copy()
    bad[24] ^= np.uint64(1)
    assert not o.verify(zz[: 4 * 9], bad)
    g.close()


def test_batched_128_msms_seed4(ctx, orc):
    rng = np.random.default_rng([4, 6])
    count = 128
    sizes = [int(2 ** (10 + 10 * u)) for u in rng.random(count)]
    nmax = 1 << 20
    s, k = orc.gen_fr(4, 0, 2 * nmax), orc.gen_fr(4, 1, nmax)
    offs = [int(o) for o in rng.integers(0, nmax, size=count)]
    ds, db = ctx.upload(s), _gpu_bases(ctx, k, nmax, 1)
    outs = ctx.alloc(96 * count)
    jobs = [(ds.ptr + 32 * o, db.ptr, m, outs.ptr + 96 * i) for i, (o, m) in enumerate(zip(offs, sizes))]
    want = np.concatenate([orc.msm_closed_form(s[4 * o: 4 * (o + m)], k[: 4 * m], 1) for o, m in zip(offs, sizes)])
    for rep in range(3):        # capture, then two replays of the graph
        outs.zero()
        ctx.msm_batch(1, jobs, graph=True)
        got = outs.download(np.uint64, 12 * count)
        bad = [i for i in range(count) if not np.array_equal(got[12 * i: 12 * i + 12], want[12 * i: 12 * i + 12])]
        assert not bad, (rep, bad[:8], [sizes[i] for i in bad[:8]])
    for b in (ds, db, outs):
        b.free()


def test_batched_128_ntts_seed4(ctx, orc):
    rng = np.random.default_rng([4, 6])
    rng.random(128), rng.integers(0, 1 << 20, size=128)     # the MSM draws come first on this stream
    logs = [int(v) for v in rng.integers(10, 21, size=128)]
    tot = sum(1 << lg for lg in logs)
    x = orc.gen_fr(4, 8, tot)
    offs, base = [], 0
    for lg in logs:
        offs.append(base)
        base += 1 << lg
    d_a, d_c = ctx.upload(x), ctx.upload(x)
    jobs = [(d_a.ptr + 32 * o, 1 << lg, 0) for lg, o in zip(logs, offs)]
    ctx.ntt_batch(jobs, graph=True)
    for lg, o in zip(logs, offs):
        ctx.ntt(d_c.ptr + 32 * o, 1 << lg)
    y = d_a.download(np.uint64, 4 * tot)
    assert np.array_equal(y, d_c.download(np.uint64, 4 * tot))
    checked = 0
    for lg, o in zip(logs, offs):
        if lg <= 12:
            n = 1 << lg
            tab = orc.ntt_setup(n)
            assert np.array_equal(y[4 * o: 4 * (o + n)], orc.ntt(x[4 * o: 4 * (o + n)], tab[: 4 * n])), lg
            checked += 1
    assert checked > 0
    n = 1 << 18
    x18 = orc.gen_fr(4, 9, n)
    d18 = ctx.upload(x18)
    ctx.ntt_batch([(d18.ptr, n, 0)], graph=True)
    assert np.array_equal(d18.download(np.uint64, 4 * n), orc.ntt(x18, orc.ntt_setup(n)[: 4 * n]))
    ctx.ntt_batch([(d_a.ptr + 32 * o, 1 << lg, 1) for lg, o in zip(logs, offs)], graph=True)
    assert np.array_equal(d_a.download(np.uint64, 4 * tot), x)
    for b in (d_a, d_c, d18):
        b.free()


def test_served_mix_2e12_to_2e20_seed5(orc):
    import liblightzk_amd as z
    logs, weights = (12, 16, 18, 20), (40, 30, 20, 10)
    srv = z.ProofServer(0, max_batch=64, mode="continuous")
    cids = {lg: srv.add_circuit(40 + lg, lg, 8) for lg in logs}
    oracles = {lg: orc.Groth16(seed=40 + lg, log_m=lg, num_public=8) for lg in logs}
    rng = np.random.default_rng(5)
    p = np.array(weights, dtype=float) / sum(weights)
    kinds = [logs[int(i)] for i in rng.choice(len(logs), size=48, p=p)]
    for lg in logs:             # every size at least once
        if lg not in kinds:
            kinds.append(lg)
    reqs = [(lg, srv.witness(cids[lg], 1000 + i), z.Groth16.rs(1000 + i)) for i, lg in enumerate(kinds)]
    with srv:
        srv.start()
        ids = [srv.submit(cids[lg], zz, rs) for lg, zz, rs in reqs]
        verified = set()
        for (lg, zz, rs), rid in zip(reqs, ids):
            st, proof, _ = srv.wait(rid, 600)
            assert st == 1, (lg, st, srv.last_error())
            assert np.array_equal(proof, oracles[lg].prove_trapdoor(zz, rs)), lg
            if lg not in verified:
                assert np.array_equal(zz, oracles[lg].witness(1000 + kinds.index(lg)))
                assert oracles[lg].verify(zz[: 4 * 9], proof), lg
                verified.add(lg)
        assert verified == set(logs)
        assert srv.stats()["max_inflight"] > 1
