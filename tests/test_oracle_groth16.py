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


# ---- the reference's MatMul statement (xpu-spmat.cuh:33-65, groth16.cuh:11-46)

# (p, k, N, M) of the reference's input_scale presets, include/protocol/groth16.cuh:39-46
REF_PRESETS = [(39, 4563, 63882, 65536), (63, 11907, 261954, 262144), (100, 30000, 1030000, 1048576)]


@pytest.mark.parametrize("p,k,N,M", REF_PRESETS)
def test_matmul_sizes_equal_reference_presets(orc, p, k, N, M):
    s = orc.matmul_sizes(p)
    assert (s["num_public"], s["num_vars"] - 1, s["M"]) == (k, N, M)
    # the reference's allocation sizes for val/col_idx (xpu-spmat.cuh:39-40)
    assert (s["nnzA"], s["nnzB"], s["nnzC"]) == (2 * p ** 3, p ** 3 + p ** 2, p ** 3 + p ** 2)
    assert s["num_gates"] == p ** 3 + p ** 2


def test_matmul_witness_is_the_matrix_product(orc):
    p = 5
    z = orc.matmul_witness(p, 9)
    v = [pm.from_mont(x, pm.R) for x in pm.elems(z)]
    p2 = p * p
    X = [v[1 + i * p: 1 + (i + 1) * p] for i in range(p)]
    Y = [v[1 + p2 + i * p: 1 + p2 + (i + 1) * p] for i in range(p)]
    Q = [v[1 + 2 * p2 + i * p: 1 + 2 * p2 + (i + 1) * p] for i in range(p)]
    assert v[0] == 1
    for a in range(p):
        for b in range(p):
            assert Q[a][b] == sum(X[a][t] * Y[t][b] for t in range(p)) % pm.R
    assert orc.matmul_satisfied(p, z)
    bad = z.copy()
    bad[4 * (1 + 2 * p2)] ^= np.uint64(1)       # Q[0][0] no longer X Y
    assert not orc.matmul_satisfied(p, bad)
    bad = z.copy()
    bad[4 * (1 + 3 * p2 + 7)] ^= np.uint64(1)   # one partial product
    assert not orc.matmul_satisfied(p, bad)


def test_matmul_csr_dense_semantics(orc):
    """Every row of the dense A, B, C over a random assignment: rows < p^3
    select X[i2][i1], Y[i1][i3], t_i; rows p^3 + i sum the t_{k p^2 + i},
    select z_0 and Q_i; padding rows are empty."""
    p = 3
    s = orc.matmul_sizes(p)
    p2, p3 = p * p, p ** 3
    dense = []
    for which in range(3):
        rp, col, val = orc.matmul_csr(p, which)
        ends = list(rp[1:]) + [col.size]
        rows = [[0] * s["num_vars"] for _ in range(s["M"])]
        for r in range(s["M"]):
            for j in range(int(rp[r]), int(ends[r])):
                rows[r][int(col[j])] += pm.from_mont(pm.limbs_to_int(val[4 * j: 4 * j + 4]), pm.R)
        dense.append(rows)
    A, B, C = dense
    for i in range(p3):
        i1, i2, i3 = i // p2, (i // p) % p, i % p
        assert [c for c, x in enumerate(A[i]) if x] == [1 + i2 * p + i1] and A[i][1 + i2 * p + i1] == 1
        assert [c for c, x in enumerate(B[i]) if x] == [1 + p2 + i1 * p + i3]
        assert [c for c, x in enumerate(C[i]) if x] == [1 + 3 * p2 + i]
    for i in range(p2):
        r = p3 + i
        assert [c for c, x in enumerate(A[r]) if x] == sorted(1 + 3 * p2 + k * p2 + i for k in range(p))
        assert [c for c, x in enumerate(B[r]) if x] == [0]
        assert [c for c, x in enumerate(C[r]) if x] == [1 + 2 * p2 + i]
    for r in range(p3 + p2, s["M"]):
        assert not any(A[r]) and not any(B[r]) and not any(C[r])


def test_matmul_proof_verifies(orc):
    o = orc.Groth16(3, matmul_p=3)
    z = o.witness(21)
    rs = o.rs(21)
    p = o.prove(z, rs)
    assert np.array_equal(p, o.prove_trapdoor(z, rs))
    npub = o.sizes["num_public"]
    assert o.verify(z[: 4 * (npub + 1)], p)
    wrong = z[: 4 * (npub + 1)].copy()
    wrong[4 * npub] ^= np.uint64(1)     # a different product matrix entry
    assert not o.verify(wrong, p)
