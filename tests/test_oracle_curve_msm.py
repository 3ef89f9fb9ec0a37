"""Pins the CPU oracle's group law and MSM (test/msm-check.cu:43-57 semantics).

Checks against the Python-integer affine model (tests/pymodel.py) and against
closed forms: bases P_i = k_i G make sum s_i P_i = (sum s_i k_i mod r) G, and
the special cases all-zero scalars, all-one scalars, repeated bases (doubling
inside a bucket), infinity bases and P + (-P).
"""
import numpy as np
import pytest

import pymodel as pm


def test_generators_on_curve_and_order(orc, golden):
    g1 = orc.g1_generator()
    g2 = orc.g2_generator()
    assert orc.lib().orc_g1_on_curve(orc.ptr(g1)) == 1
    assert orc.lib().orc_g2_on_curve(orc.ptr(g2)) == 1
    assert pm.g1_jac_from_buf(g1) == pm.G1_GEN
    assert pm.g2_jac_from_buf(g2) == pm.G2_GEN
    x0, x1, y0, y1 = (int(v, 16) for v in golden["g2_one"])
    assert pm.G2_GEN == ((x0, x1), (y0, y1))
    assert pm.G1C.on_curve(pm.G1_GEN) and pm.G2C.on_curve(pm.G2_GEN)
    assert pm.G1C.smul(pm.R, pm.G1_GEN) is None
    assert pm.G2C.smul(pm.R, pm.G2_GEN) is None


def test_mul_gen_matches_model(orc):
    k = orc.gen_fr(5, 1, 6)
    ks = [pm.from_mont(v, pm.R) for v in pm.elems(k)]
    p1 = orc.g1_mul_gen(k)
    p2 = orc.g2_mul_gen(k)
    for i, kc in enumerate(ks):
        assert pm.g1_aff_from_buf(p1[8 * i: 8 * i + 8]) == pm.G1C.smul(kc, pm.G1_GEN)
        assert pm.g2_aff_from_buf(p2[16 * i: 16 * i + 16]) == pm.G2C.smul(kc, pm.G2_GEN)


def _jac_from_aff(aff, width):
    """affine buffer -> Jacobian buffer with Z = R mod p (Montgomery one)."""
    one = pm.int_to_limbs(pm.MONT % pm.P)
    n = aff.size // width
    parts = []
    for i in range(n):
        a = aff[width * i: width * (i + 1)]
        if width == 8:
            parts += [a, one]
        else:
            parts += [a, one, np.zeros(4, dtype=np.uint64)]
    return np.concatenate(parts)


def test_group_ops_match_model(orc):
    k = orc.gen_fr(6, 1, 8)
    a1 = orc.g1_mul_gen(k)
    j1 = _jac_from_aff(a1, 8)
    b1 = np.roll(j1, 12)
    add = orc.field_op  # silence linters
    out_add = np.zeros_like(j1)
    out_dbl = np.zeros_like(j1)
    orc.lib().orc_g1_op(0, orc.ptr(j1), orc.ptr(b1), orc.ptr(out_add), 8)
    orc.lib().orc_g1_op(1, orc.ptr(j1), None, orc.ptr(out_dbl), 8)
    for i in range(8):
        P = pm.g1_jac_from_buf(j1[12 * i: 12 * i + 12])
        Q = pm.g1_jac_from_buf(b1[12 * i: 12 * i + 12])
        assert pm.g1_jac_from_buf(out_add[12 * i: 12 * i + 12]) == pm.G1C.padd(P, Q)
        assert pm.g1_jac_from_buf(out_dbl[12 * i: 12 * i + 12]) == pm.G1C.padd(P, P)
    # P + P through add (doubling detection), P + (-P) = infinity
    out = np.zeros(12, dtype=np.uint64)
    orc.lib().orc_g1_op(0, orc.ptr(j1[:12].copy()), orc.ptr(j1[:12].copy()), orc.ptr(out), 1)
    assert pm.g1_jac_from_buf(out) == pm.G1C.padd(pm.g1_jac_from_buf(j1[:12]), pm.g1_jac_from_buf(j1[:12]))
    neg = np.zeros(12, dtype=np.uint64)
    orc.lib().orc_g1_op(2, orc.ptr(j1[:12].copy()), None, orc.ptr(neg), 1)
    orc.lib().orc_g1_op(0, orc.ptr(j1[:12].copy()), orc.ptr(neg), orc.ptr(out), 1)
    assert pm.g1_jac_from_buf(out) is None
    # G2
    a2 = orc.g2_mul_gen(k)
    j2 = _jac_from_aff(a2, 16)
    b2 = np.roll(j2, 24)
    o2 = np.zeros_like(j2)
    orc.lib().orc_g2_op(0, orc.ptr(j2), orc.ptr(b2), orc.ptr(o2), 8)
    for i in range(8):
        P = pm.g2_jac_from_buf(j2[24 * i: 24 * i + 24])
        Q = pm.g2_jac_from_buf(b2[24 * i: 24 * i + 24])
        assert pm.g2_jac_from_buf(o2[24 * i: 24 * i + 24]) == pm.G2C.padd(P, Q)


@pytest.mark.parametrize("n", [1, 7, 64])
def test_msm_g1_brute_force(orc, n):
    s = orc.gen_fr(21, 0, n)
    k = orc.gen_fr(21, 1, n)
    bases = orc.g1_mul_gen(k)
    want = None
    for i in range(n):
        want = pm.G1C.padd(want, pm.G1C.smul(pm.from_mont(pm.elems(s[4 * i: 4 * i + 4])[0], pm.R),
                                              pm.g1_aff_from_buf(bases[8 * i: 8 * i + 8])))
    for method in (0, 1):
        assert pm.g1_jac_from_buf(orc.msm_g1(s, bases, method)) == want
    assert pm.g1_jac_from_buf(orc.msm_closed_form(s, k, 1)) == want


def test_msm_g1_special_cases(orc):
    n = 300
    k = orc.gen_fr(22, 1, n)
    bases = orc.g1_mul_gen(k)
    zero = np.zeros(4 * n, dtype=np.uint64)
    assert pm.g1_jac_from_buf(orc.msm_g1(zero, bases)) is None
    one = np.tile(pm.int_to_limbs(pm.MONT % pm.R), n)
    want = None
    for i in range(n):
        want = pm.G1C.padd(want, pm.g1_aff_from_buf(bases[8 * i: 8 * i + 8]))
    assert pm.g1_jac_from_buf(orc.msm_g1(one, bases, 1)) == want
    # repeated bases + infinity bases + sparse scalars
    rep = np.tile(bases[:8 * 10], n // 10)
    rep[8 * 5: 8 * 6] = 0
    s = orc.gen_fr(22, 8, n, sparse=True)
    r0 = orc.msm_g1(s, rep, 0)
    r1 = orc.msm_g1(s, rep, 1)
    assert np.array_equal(r0, r1)
    want = None
    for i in range(n):
        pt = pm.g1_aff_from_buf(rep[8 * i: 8 * i + 8])
        want = pm.G1C.padd(want, pm.G1C.smul(pm.from_mont(pm.elems(s[4 * i: 4 * i + 4])[0], pm.R), pt) if pt else None)
    assert pm.g1_jac_from_buf(r1) == want


def test_msm_g2_closed_form(orc):
    n = 40
    s = orc.gen_fr(23, 0, n)
    k = orc.gen_fr(23, 1, n)
    bases = orc.g2_mul_gen(k)
    r0 = orc.msm_g2(s, bases, 0)
    r1 = orc.msm_g2(s, bases, 1)
    cf = orc.msm_closed_form(s, k, 2)
    assert np.array_equal(r0, r1) and np.array_equal(r1, cf)
    sk = sum(pm.from_mont(a, pm.R) * pm.from_mont(b, pm.R) for a, b in zip(pm.elems(s), pm.elems(k))) % pm.R
    assert pm.g2_jac_from_buf(cf) == pm.G2C.smul(sk, pm.G2_GEN)


def test_msm_g1_closed_form_larger(orc):
    n = 1 << 12
    s = orc.gen_fr(1, 0, n)
    k = orc.gen_fr(1, 1, n)
    bases = orc.g1_mul_gen(k)
    assert np.array_equal(orc.msm_g1(s, bases, 1), orc.msm_closed_form(s, k, 1))
