"""GPU NTT / coset NTT / SpMV / pointwise vs the CPU oracle, bit-exact.

NTT semantics are the reference operator's (include/operator/poly.cuh:32-73:
out[k] = factor * sum_j in[j] w^(jk) with the ntt_setup table,
include/protocol/groth16-operator.cuh:6-32), so the oracle's restatement of
that exact stage loop is the checker, with the reference's own twiddle tables
(built by the oracle's ntt_setup, whose roots are pinned to
include/host-arith/ntt-params.cuh). Sizes 2^1..2^16 cover single-pass,
two-pass and clamped-column pass plans; 2^20 and 2^22 are checked by round
trips (NTT then inverse NTT is the identity) and against the oracle on a
2^18 instance."""
import numpy as np
import pytest

import pymodel as pm

pytestmark = pytest.mark.gpu


def _coset_expected(orc, x, n, tab_fwd):
    g = pm.to_mont(5, pm.R)
    pw = [pm.to_mont(pow(5, j, pm.R), pm.R) for j in range(n)]
    scaled = orc.field_op("fr", 0, x, np.concatenate([pm.int_to_limbs(v) for v in pw]))
    return orc.ntt(scaled, tab_fwd)


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8, 10, 11, 12, 13, 14, 16])
def test_ntt_reference_table(ctx, orc, log_n):
    n = 1 << log_n
    x = orc.gen_fr(41, log_n, n)
    tab = orc.ntt_setup(n)
    fwd, inv, minv = tab[: 4 * n], tab[4 * n: 8 * n], tab[8 * n: 8 * n + 4]
    d = ctx.upload(x)
    dt_f, dt_i, dm = ctx.upload(fwd), ctx.upload(inv), ctx.upload(minv)
    ctx.ntt(d, n, dt_f)
    y = d.download(np.uint64, 4 * n)
    assert np.array_equal(y, orc.ntt(x, fwd)), log_n
    ctx.ntt(d, n, dt_i, factor=dm)
    assert np.array_equal(d.download(np.uint64, 4 * n), x), log_n


@pytest.mark.parametrize("log_n", [1, 4, 11, 12, 15])
def test_ntt_library_tables_and_coset(ctx, orc, log_n):
    n = 1 << log_n
    x = orc.gen_fr(42, log_n, n)
    tab = orc.ntt_setup(n)
    d = ctx.upload(x)
    ctx.ntt(d, n)                         # library forward table
    assert np.array_equal(d.download(np.uint64, 4 * n), orc.ntt(x, tab[: 4 * n]))
    ctx.ntt(d, n, inverse=True)           # library inverse, scaled by 1/n
    assert np.array_equal(d.download(np.uint64, 4 * n), x)
    ctx.coset_ntt(d, n)
    assert np.array_equal(d.download(np.uint64, 4 * n), _coset_expected(orc, x, n, tab[: 4 * n]))
    ctx.coset_ntt(d, n, inverse=True)
    assert np.array_equal(d.download(np.uint64, 4 * n), x)


def test_ntt_delta_and_constant(ctx, orc):
    n = 1 << 13
    one = pm.int_to_limbs(pm.MONT % pm.R)
    delta = np.zeros(4 * n, dtype=np.uint64)
    delta[:4] = one
    d = ctx.upload(delta)
    ctx.ntt(d, n)
    assert np.array_equal(d.download(np.uint64, 4 * n), np.tile(one, n))   # NTT(delta_0) = 1...1
    ctx.ntt(d, n)   # NTT(all ones) = n * delta_0
    want = np.zeros(4 * n, dtype=np.uint64)
    want[:4] = pm.int_to_limbs(pm.to_mont(n, pm.R))
    assert np.array_equal(d.download(np.uint64, 4 * n), want)


# a lone transform's threads hold 2 elements up to 2^19 and 8 from 2^20 on;
# both forms are compared with the oracle
@pytest.mark.parametrize("log_n", [18, 19, 20, 22])
def test_ntt_large(ctx, orc, log_n):
    n = 1 << log_n
    x = orc.gen_fr(43, log_n, n)
    d = ctx.upload(x)
    ctx.ntt(d, n)
    if log_n <= 20:
        tab = orc.ntt_setup(n)
        assert np.array_equal(d.download(np.uint64, 4 * n), orc.ntt(x, tab[: 4 * n]))
    ctx.ntt(d, n, inverse=True)
    assert np.array_equal(d.download(np.uint64, 4 * n), x)
    ctx.coset_ntt(d, n)
    ctx.coset_ntt(d, n, inverse=True)
    assert np.array_equal(d.download(np.uint64, 4 * n), x)


def test_ntt_rejects_non_power_of_two(ctx):
    from liblightzk_amd import HzkError
    d = ctx.alloc(32 * 12)
    with pytest.raises(HzkError):
        ctx.ntt(d, 12)


def test_spmv3_and_pointwise(ctx, orc):
    import liblightzk_amd as z
    seed, log_m, npub = 7, 10, 8
    sizes = orc.circuit_sizes(seed, log_m, npub)
    M, nv = sizes["M"], sizes["num_vars"]
    zz = orc.circuit_witness(seed, log_m, npub, 77)
    dz = ctx.upload(zz)
    mats, keep = [], []
    for which in range(3):
        rp, col, val = orc.circuit_csr(seed, log_m, npub, which)
        dv, dc, dr = ctx.upload(val), ctx.upload(col), ctx.upload(rp)
        keep += [dv, dc, dr]
        mats.append(z.Csr(dv.ptr, dc.ptr, dr.ptr, M, nv, col.size))
    outs = [ctx.alloc(32 * M) for _ in range(3)]
    z.batch_sparse_matrix_vector_multiplication(ctx, *mats, dz, *outs)
    got = [o.download(np.uint64, 4 * M) for o in outs]
    for which in range(3):
        rp, col, val = orc.circuit_csr(seed, log_m, npub, which)
        assert np.array_equal(got[which], orc.spmv(M, col.size, val, col, rp, zz)), which
    # A o B - C vanishes on a satisfying witness
    z.polynomial_A_times_B_minus_C(ctx, M, outs[0], outs[1], outs[2])
    assert not outs[0].download(np.uint64, 4 * M).any()


@pytest.mark.parametrize("deg", [1, 7, 1000, 1 << 16])
def test_pointwise_abc_random(ctx, orc, deg):
    """A <- A o B - C (poly.cuh:75-86) on random, unrelated A, B, C: the
    oracle's Fr product and difference, element by element; B and C unchanged.
    Boundary values (0, 1, r - 1) lead the vectors."""
    import liblightzk_amd as z
    a, b, c = (orc.gen_fr(90 + i, 0, deg) for i in range(3))
    edge = [pm.int_to_limbs(v) for v in (0, pm.MONT % pm.R, pm.R - 1)]
    for j, e in enumerate(edge[: min(3, deg)]):
        a[4 * j: 4 * j + 4] = e
        b[4 * (deg - 1 - j): 4 * (deg - j)] = e
    da, db, dc = ctx.upload(a), ctx.upload(b), ctx.upload(c)
    z.polynomial_A_times_B_minus_C(ctx, deg, da, db, dc)
    want = orc.field_op("fr", 2, orc.field_op("fr", 0, a, b), c)
    assert np.array_equal(da.download(np.uint64, 4 * deg), want)
    assert np.array_equal(db.download(np.uint64, 4 * deg), b)
    assert np.array_equal(dc.download(np.uint64, 4 * deg), c)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("small", [False, True])
def test_ntt_batch_uneven_sizes_and_kinds(ctx, orc, graph, small):
    """Batched transforms (one launch per pass round over all transforms,
    optionally one hipGraph) equal the same transforms run one at a time, which
    are pinned to the oracle above; the forward NTTs are also checked directly.
    A batch of 2^20 elements or more runs 8-element threads; a smaller one
    (small) runs 2-element threads and the blocked permutation with its scale
    and coset factors."""
    rng = np.random.default_rng(9)
    if small:
        logs = [int(v) for v in rng.integers(1, 15, size=30)] + [0, 8, 11, 12]
        assert sum(1 << lg for lg in logs) < 1 << 20
    else:
        logs = [int(v) for v in rng.integers(1, 17, size=30)] + [0, 11, 12, 20]
    kinds = [int(v) for v in rng.integers(0, 4, size=len(logs))]
    xs = [orc.gen_fr(50 + i, 0, 1 << lg) for i, lg in enumerate(logs)]
    for rep in range(2):
        bufs = [ctx.upload(x) for x in xs]
        ctx.ntt_batch([(b, 1 << lg, kd) for b, lg, kd in zip(bufs, logs, kinds)], graph=graph)
        for i, (lg, kd) in enumerate(zip(logs, kinds)):
            n = 1 << lg
            ref = ctx.upload(xs[i])
            if kd == 0:
                ctx.ntt(ref, n)
            elif kd == 1:
                ctx.ntt(ref, n, inverse=True)
            else:
                ctx.coset_ntt(ref, n, inverse=(kd == 3))
            got = bufs[i].download(np.uint64, 4 * n)
            assert np.array_equal(got, ref.download(np.uint64, 4 * n)), (rep, i, lg, kd)
            if kd == 0 and lg <= 12:
                assert np.array_equal(got, orc.ntt(xs[i], orc.ntt_setup(n)[: 4 * n])), (i, lg)
        xs = [orc.gen_fr(80 + i, 0, 1 << lg) for i, lg in enumerate(logs)]
