"""Pins the CPU oracle's field arithmetic and constants.

Pinned against (a) the reference's own constant tables (tests/golden,
extracted from include/host-arith/ntt-params.cuh and alt_bn128_init.cu) and
(b) an independent Python-integer model (tests/pymodel.py). Edge values
(0, 1, p-1, 2^254-ish) are included because a dropped carry or a missing
final subtraction shows up exactly there.
"""
import numpy as np
import pytest

import pymodel as pm

OPS = {"mul": 0, "add": 1, "sub": 2, "neg": 3, "sqr": 4, "inv": 5, "to_canon": 6, "from_canon": 7}


def _edge_and_random(orc, kind, n):
    m = pm.P if kind == "fq" else pm.R
    edge = [0, 1, 2, m - 1, m - 2, (m - 1) // 2, (1 << 253), m >> 1, pm.MONT % m]
    rnd = pm.elems(orc.gen_fq(11, 0, n) if kind == "fq" else orc.gen_fr(11, 0, n))
    vals = edge + rnd
    return vals, np.concatenate([pm.int_to_limbs(v) for v in vals])


@pytest.mark.parametrize("kind", ["fq", "fr"])
def test_field_ops_match_int_model(orc, kind):
    m = pm.P if kind == "fq" else pm.R
    a_vals, a = _edge_and_random(orc, kind, 200)
    b_vals = list(reversed(a_vals))
    b = np.concatenate([pm.int_to_limbs(v) for v in b_vals])
    rinv = pow(pm.MONT, -1, m)
    got = {op: pm.elems(orc.field_op(kind, code, a, b)) for op, code in OPS.items()}
    for i, (x, y) in enumerate(zip(a_vals, b_vals)):
        assert got["mul"][i] == x * y * rinv % m
        assert got["add"][i] == (x + y) % m
        assert got["sub"][i] == (x - y) % m
        assert got["neg"][i] == (-x) % m
        assert got["sqr"][i] == x * x * rinv % m
        # Montgomery inverse: (xR^-1)^-1 R = x^-1 R^2
        assert got["inv"][i] == (0 if x == 0 else pow(x, -1, m) * pm.MONT * pm.MONT % m)
        assert got["to_canon"][i] == x * rinv % m
        assert got["from_canon"][i] == x * pm.MONT % m


def test_reference_constants(golden):
    assert int(golden["modulus_q"], 16) == pm.P
    assert int(golden["modulus_r"], 16) == pm.R
    for m, r2, r3, inv, dev in ((pm.P, "fq_Rsquared", "fq_Rcubed", "fq_inv", "device_M0_fq"),
                                (pm.R, "fr_Rsquared", "fr_Rcubed", "fr_inv", "device_m0_fr")):
        assert int(golden[r2], 16) == pm.MONT ** 2 % m
        assert int(golden[r3], 16) == pm.MONT ** 3 % m
        assert int(golden[inv], 16) == (-pow(m, -1, 1 << 64)) % (1 << 64)
        assert int(golden[dev], 16) == (-pow(m, -1, 1 << 32)) % (1 << 32)
    assert int(golden["ate_loop_count"], 16) == 6 * 4965661367192848881 + 2
    assert int(golden["final_exponent"], 16) == (pm.P ** 12 - 1) // pm.R


def test_root_tables_match_reference(orc, golden):
    """Oracle roots (5^((r-1)/2^k)) equal the reference tables byte-for-byte."""
    fwd, inv, scl = (golden["ntt_forward_root_of_unity_mont"], golden["ntt_inverse_root_of_unity_mont"],
                     golden["ntt_domain_scale_inverse_mont"])
    for k in range(29):
        w = pm.limbs_to_int(orc.root_of_unity(k))
        wi = pm.limbs_to_int(orc.root_of_unity(k, inverse=True))
        assert w == int(fwd[k], 16), k
        assert wi == int(inv[k], 16), k
        # independent checks: primitive 2^k-th root, inverse, scale 2^-k
        wc = pm.from_mont(w, pm.R)
        assert pow(wc, 1 << k, pm.R) == 1
        if k:
            assert pow(wc, 1 << (k - 1), pm.R) == pm.R - 1
        assert pm.from_mont(int(scl[k], 16), pm.R) == pow(pow(2, k, pm.R), -1, pm.R)
    assert pm.from_mont(int(fwd[28], 16), pm.R) == int(golden["fr_root_of_unity"], 16)


def test_fq2_ops(orc):
    n = 64
    a = orc.gen_fq(12, 0, 2 * n)
    b = orc.gen_fq(12, 1, 2 * n)
    got = {op: orc.field_op("fq2", code, a, b) for op, code in (("mul", 0), ("add", 1), ("sub", 2), ("sqr", 4), ("inv", 5))}
    A = [pm.from_mont(v, pm.P) for v in pm.elems(a)]
    B = [pm.from_mont(v, pm.P) for v in pm.elems(b)]
    for i in range(n):
        x, y = (A[2 * i], A[2 * i + 1]), (B[2 * i], B[2 * i + 1])

        def g(op):
            v = pm.elems(got[op][8 * i: 8 * i + 8])
            return (pm.from_mont(v[0], pm.P), pm.from_mont(v[1], pm.P))
        assert g("mul") == pm.f2_mul(x, y)
        assert g("add") == pm.f2_add(x, y)
        assert g("sub") == pm.f2_sub(x, y)
        assert g("sqr") == pm.f2_mul(x, x)
        assert g("inv") == pm.f2_inv(x)


def test_rng_is_deterministic_and_in_range(orc):
    a = orc.gen_fr(1, 0, 1000)
    b = orc.gen_fr(1, 0, 1000)
    c = orc.gen_fr(1, 1, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    vals = pm.elems(a)
    assert all(v < pm.R for v in vals)
    assert len(set(vals)) == 1000
    s = pm.elems(orc.gen_fr(3, 8, 4000, sparse=True))
    one_m = pm.MONT % pm.R
    frac01 = sum(1 for v in s if v in (0, one_m)) / len(s)
    assert 0.65 < frac01 < 0.75


def test_rng_splitmix_xoshiro_known_answer(orc):
    """First outputs of the seeded stream, computed here from the published
    SplitMix64 and xoshiro256** definitions."""
    M = (1 << 64) - 1

    def splitmix(x):
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return x, z ^ (z >> 31)

    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & M

    seed, stream = 7, 3
    x = ((seed << 32) ^ stream) & M
    s = []
    for _ in range(4):
        x, z = splitmix(x)
        s.append(z)
    outs = []
    for _ in range(8):
        res = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        outs.append(res)
    assert [int(v) for v in orc.rng_u64(seed, stream, 8)] == outs
