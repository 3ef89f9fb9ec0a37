"""Pins the oracle pairing and Groth16 (the reference has neither: groth16.cuh
uses random keys and no verifier, alt_bn128_init.cu:83-112 leaves Fq6/Fq12
commented out).

Pairing: bilinearity e(aP, bQ) = e(P, Q)^(ab), non-degeneracy, e(P, Q)^r = 1,
and the Frobenius twist constants equal the reference's
alt_bn128_twist_mul_by_q_X/Y (alt_bn128_init.cu:121-124).
Groth16: R1CS satisfaction, QAP identity at tau, H by Schwartz-Zippel,
honest proof == trapdoor closed form, verifier accepts honest proofs and
rejects tampered proofs and wrong public inputs.
"""
import numpy as np
import pytest

import pymodel as pm


def test_twist_frobenius_constants(orc, golden):
    out = np.zeros(16, dtype=np.uint64)
    orc.lib().orc_twist_frobenius_consts(orc.ptr(out))
    v = [pm.from_mont(t, pm.P) for t in pm.elems(out)]
    assert (v[0], v[1]) == tuple(int(x, 16) for x in golden["twist_mul_by_q_X"])
    assert (v[2], v[3]) == tuple(int(x, 16) for x in golden["twist_mul_by_q_Y"])


def _fq12(buf):
    return tuple(pm.from_mont(v, pm.P) for v in pm.elems(buf))


def test_pairing_bilinear_nondegenerate(orc):
    k = orc.gen_fr(41, 0, 2)
    a, b = (pm.from_mont(v, pm.R) for v in pm.elems(k))
    P = pm.g1_to_aff_buf(pm.G1_GEN)
    Q = pm.g2_to_aff_buf(pm.G2_GEN)
    aP = pm.g1_to_aff_buf(pm.G1C.smul(a, pm.G1_GEN))
    bQ = pm.g2_to_aff_buf(pm.G2C.smul(b, pm.G2_GEN))
    e = orc.pairing(P, Q)
    one = tuple([1] + [0] * 11)
    assert _fq12(e) != one
    # e(aP, bQ) == e(P, Q)^(ab) checked as e(aP, bQ) * e(-abP, Q) == 1
    ab = a * b % pm.R
    ps = np.concatenate([aP, pm.g1_to_aff_buf(pm.G1C.neg(pm.G1C.smul(ab, pm.G1_GEN)))])
    qs = np.concatenate([bQ, Q])
    assert orc.pairing_product_is_one(ps, qs)
    # and a wrong exponent is rejected
    ps_bad = np.concatenate([aP, pm.g1_to_aff_buf(pm.G1C.neg(pm.G1C.smul(ab + 1, pm.G1_GEN)))])
    assert not orc.pairing_product_is_one(ps_bad, qs)
    # e(P, Q)^r == 1: e(rP, Q) with rP = infinity gives 1 and e(P,Q)*e(-P,Q) == 1
    assert orc.pairing_product_is_one(np.concatenate([P, pm.g1_to_aff_buf(pm.G1C.neg(pm.G1_GEN))]),
                                      np.concatenate([Q, Q]))
    assert _fq12(orc.pairing(np.zeros(8, dtype=np.uint64), Q)) == one


@pytest.fixture(scope="module")
def g16(orc):
    return orc.Groth16(seed=2, log_m=7, num_public=8)


def test_circuit_and_witness(orc, g16):
    s = g16.sizes
    assert s["M"] == 128 and s["num_gates"] == 128 - 9 and s["num_vars"] == 1 + 64 + 119
    z = g16.witness(100)
    assert orc.r1cs_satisfied(2, 7, 8, z)
    z2 = z.copy()
    z2[4 * 70] ^= np.uint64(1)
    assert not orc.r1cs_satisfied(2, 7, 8, z2)
    rp, col, val = orc.circuit_csr(2, 7, 8, 0)
    assert rp.size == 128 and int(rp[0]) == 0 and np.all(np.diff(rp.astype(np.int64)) >= 1)


def test_qap_identity_and_h(orc, g16):
    z = g16.witness(101)
    zs = [pm.from_mont(v, pm.R) for v in pm.elems(z)]
    At, Bt, Ct = ([pm.from_mont(v, pm.R) for v in pm.elems(g16.key_scalars(w))] for w in (0, 1, 2))
    tau = pm.from_mont(pm.elems(g16.key_scalars(6))[0], pm.R)
    M = g16.sizes["M"]
    a = sum(x * y for x, y in zip(zs, At)) % pm.R
    b = sum(x * y for x, y in zip(zs, Bt)) % pm.R
    c = sum(x * y for x, y in zip(zs, Ct)) % pm.R
    h = [pm.from_mont(v, pm.R) for v in pm.elems(g16.h_coeffs(z))]
    assert h[M - 1] == 0
    ht = sum(hj * pow(tau, j, pm.R) for j, hj in enumerate(h)) % pm.R
    assert (a * b - c) % pm.R == ht * (pow(tau, M, pm.R) - 1) % pm.R


def test_key_points_match_scalars(orc, g16):
    At = g16.key_scalars(0)
    A1 = g16.key_points(0)
    for i in (0, 1, 9, 50, 120):
        kc = pm.from_mont(pm.elems(At[4 * i: 4 * i + 4])[0], pm.R)
        assert pm.g1_aff_from_buf(A1[8 * i: 8 * i + 8]) == pm.G1C.smul(kc, pm.G1_GEN)


def test_prove_verify(orc, g16):
    z = g16.witness(102)
    rs = g16.rs(102)
    p_honest = g16.prove(z, rs)
    p_trap = g16.prove_trapdoor(z, rs)
    assert np.array_equal(p_honest, p_trap)
    public = z[: 4 * (g16.num_public + 1)]
    assert g16.verify(public, p_honest)
    bad = p_honest.copy()
    bad[24:32] = pm.g1_to_aff_buf(pm.G1C.padd(pm.g1_aff_from_buf(bad[24:32]), pm.G1_GEN))
    assert not g16.verify(public, bad)
    wrong_pub = public.copy()
    wrong_pub[4:8] = pm.int_to_limbs(pm.to_mont(12345, pm.R))
    assert not g16.verify(wrong_pub, p_honest)
