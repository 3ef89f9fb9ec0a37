"""The oracle's restatement against the reference's own compiled host
arithmetic (oracle/_ref/libref_hostarith.so, built by oracle/ref/Makefile from
the reference sources; skipped where that library is absent).

Same seeded inputs into both; results must be byte-identical: Fq/Fr/Fq2
operations, G1/G2 addition (incl. doubling and P + (-P)), scalar_mul
(curve_utils.cuh:29-50) and the naive MSM of test/msm-check.cu:46-55.
"""
import ctypes
import os

import numpy as np
import pytest

import pymodel as pm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_hostarith.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference tree)")


@pytest.fixture(scope="module")
def ref():
    L = ctypes.CDLL(REF_SO)
    P = ctypes.POINTER(ctypes.c_uint64)
    for name in ("ref_fq_op", "ref_fr_op", "ref_fq2_op"):
        getattr(L, name).argtypes = [ctypes.c_int, P, P, P, ctypes.c_size_t]
    for name in ("ref_g1_scalar_mul", "ref_g2_scalar_mul", "ref_g1_add", "ref_g2_add"):
        getattr(L, name).argtypes = [P, P, P, ctypes.c_size_t]
    for name in ("ref_msm_g1", "ref_msm_g2"):
        getattr(L, name).argtypes = [P, P, ctypes.c_size_t, P]
    L.ref_init()
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


@pytest.mark.parametrize("kind", ["fq", "fr", "fq2"])
def test_field_ops_equal_reference(orc, ref, kind):
    n = 3000
    gen = orc.gen_fr if kind == "fr" else orc.gen_fq
    width = 8 if kind == "fq2" else 4
    a = gen(31, 0, n * width // 4)
    b = gen(31, 1, n * width // 4)
    a[:width] = 0
    b[width: 2 * width] = a[width: 2 * width]
    fn = {"fq": ref.ref_fq_op, "fr": ref.ref_fr_op, "fq2": ref.ref_fq2_op}[kind]
    for op in range(6):
        want = np.zeros_like(a)
        fn(op, _p(a), _p(b), _p(want), n)
        got = orc.field_op(kind, op, a, b)
        assert np.array_equal(got, want), (kind, op)


def test_group_ops_equal_reference(orc, ref):
    n = 64
    k = orc.gen_fr(32, 0, n)
    s = orc.gen_fr(32, 1, n)
    p1 = orc.g1_mul_gen(k)
    p2 = orc.g2_mul_gen(k)
    # scalar_mul
    w1 = np.zeros(12 * n, dtype=np.uint64)
    w2 = np.zeros(24 * n, dtype=np.uint64)
    ref.ref_g1_scalar_mul(_p(s), _p(p1), _p(w1), n)
    ref.ref_g2_scalar_mul(_p(s), _p(p2), _p(w2), n)
    for i in range(n):
        g1 = np.zeros(12, dtype=np.uint64)
        orc.lib().orc_g1_scalar_mul(orc.ptr(s[4 * i: 4 * i + 4].copy()), orc.ptr(p1[8 * i: 8 * i + 8].copy()), orc.ptr(g1))
        assert np.array_equal(g1, w1[12 * i: 12 * i + 12]), i
        g2 = np.zeros(24, dtype=np.uint64)
        orc.lib().orc_g2_scalar_mul(orc.ptr(s[4 * i: 4 * i + 4].copy()), orc.ptr(p2[16 * i: 16 * i + 16].copy()), orc.ptr(g2))
        assert np.array_equal(g2, w2[24 * i: 24 * i + 24]), i
    # add: generic, doubling (same point), inverse (P + -P), infinity operand
    q1 = np.roll(p1, 8)
    q1[8:16] = p1[8:16]
    q1[16:24] = pm.g1_to_aff_buf(pm.G1C.neg(pm.g1_aff_from_buf(p1[16:24])))
    q1[24:32] = 0
    wa = np.zeros(12 * n, dtype=np.uint64)
    ref.ref_g1_add(_p(p1), _p(q1), _p(wa), n)
    for i in range(n):
        A = pm.g1_aff_from_buf(p1[8 * i: 8 * i + 8])
        B = pm.g1_aff_from_buf(q1[8 * i: 8 * i + 8])
        assert pm.g1_jac_from_buf(wa[12 * i: 12 * i + 12]) == pm.G1C.padd(A, B), i


def test_naive_msm_equals_reference(orc, ref):
    n = 2000
    s = orc.gen_fr(33, 0, n, sparse=True)
    k = orc.gen_fr(33, 1, n)
    b1 = orc.g1_mul_gen(k)
    b2 = orc.g2_mul_gen(k[: 4 * 300].copy())
    w1 = np.zeros(12, dtype=np.uint64)
    ref.ref_msm_g1(_p(s), _p(b1), n, _p(w1))
    assert np.array_equal(orc.msm_g1(s, b1, 0), w1)
    assert np.array_equal(orc.msm_g1(s, b1, 1), w1)
    w2 = np.zeros(24, dtype=np.uint64)
    s2 = s[: 4 * 300].copy()
    ref.ref_msm_g2(_p(s2), _p(b2), 300, _p(w2))
    assert np.array_equal(orc.msm_g2(s2, b2, 0), w2)
