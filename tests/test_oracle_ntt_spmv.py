"""Pins the oracle's NTT operator (include/operator/poly.cuh:32-73), twiddle
table (include/protocol/groth16-operator.cuh:6-32) and CSR SpMV (poly.cuh:5-30).

NTT: DFT by definition for small n, round trip, NTT(delta_0) = all ones, and
the convolution theorem. SpMV: dense brute force in Python integers.
"""
import numpy as np
import pytest

import pymodel as pm


def _mont_list(vals):
    return np.concatenate([pm.int_to_limbs(pm.to_mont(v, pm.R)) for v in vals])


def _canon(buf):
    return [pm.from_mont(v, pm.R) for v in pm.elems(buf)]


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 7])
def test_ntt_matches_dft(orc, log_n):
    n = 1 << log_n
    x = orc.gen_fr(31, log_n, n)
    M = max(n, 2)
    tab = orc.ntt_setup(n)
    fwd = tab[:4 * n]
    out = orc.ntt(x, fwd)
    w = pm.from_mont(pm.elems(fwd[4:8])[0], pm.R) if n > 1 else 1
    assert pow(w, n, pm.R) == 1
    assert _canon(out) == pm.dft(_canon(x), w, pm.R)
    # inverse with the table's M^-1 factor round-trips
    inv = tab[4 * n: 8 * n]
    minv = tab[8 * n: 8 * n + 4]
    back = orc.ntt(out, inv, minv)
    assert np.array_equal(back, x)


def test_ntt_setup_layout(orc):
    M = 1 << 4
    t = orc.ntt_setup(M)
    T = _canon(t)
    wM = T[1]
    w2M = T[2 * M + 2]
    assert pow(wM, M, pm.R) == 1 and pow(wM, M // 2, pm.R) == pm.R - 1
    assert pow(w2M, 2 * M, pm.R) == 1 and pow(w2M, M, pm.R) == pm.R - 1
    for i in range(M):
        assert T[i] == pow(wM, i, pm.R)
        assert T[M + i] == pow(wM, -i, pm.R)
    assert T[2 * M] == pow(M, -1, pm.R)
    for i in range(2 * M):
        assert T[2 * M + 1 + i] == pow(w2M, i, pm.R)
        assert T[4 * M + 1 + i] == pow(w2M, -i, pm.R)
    # the reference stores (4M)^-1 at 6M+1 although its layout comment says (2M)^-1
    assert T[6 * M + 1] == pow(4 * M, -1, pm.R)


def test_ntt_delta_and_convolution(orc):
    n = 1 << 8
    tab = orc.ntt_setup(n)
    fwd, inv, minv = tab[:4 * n], tab[4 * n: 8 * n], tab[8 * n: 8 * n + 4]
    delta = _mont_list([1] + [0] * (n - 1))
    assert _canon(orc.ntt(delta, fwd)) == [1] * n
    a = [v % pm.R for v in range(1, n // 2 + 1)] + [0] * (n // 2)
    b = [(7 * v * v + 3) % pm.R for v in range(n // 2)] + [0] * (n // 2)
    A = orc.ntt(_mont_list(a), fwd)
    B = orc.ntt(_mont_list(b), fwd)
    prod = orc.field_op("fr", 0, A, B)
    c = _canon(orc.ntt(prod, inv, minv))
    want = [0] * n
    for i, ai in enumerate(a):
        for j, bj in enumerate(b):
            if ai and bj:
                want[(i + j) % n] = (want[(i + j) % n] + ai * bj) % pm.R
    assert c == want


def test_spmv_dense_brute_force(orc):
    rng = np.random.default_rng(5)
    rows, cols = 37, 23
    dense = [[0] * cols for _ in range(rows)]
    row_ptr, col_idx, vals = [], [], []
    for i in range(rows):
        row_ptr.append(len(col_idx))
        for _ in range(int(rng.integers(0, 4))):
            j = int(rng.integers(0, cols))
            v = int(rng.integers(1, 1 << 62)) * 977 % pm.R
            col_idx.append(j); vals.append(v)
            dense[i][j] = (dense[i][j] + v) % pm.R
    z = [int(rng.integers(0, 1 << 62)) ** 3 % pm.R for _ in range(cols)]
    out = orc.spmv(rows, len(col_idx), _mont_list(vals), np.array(col_idx, dtype=np.uint64),
                   np.array(row_ptr, dtype=np.uint64), _mont_list(z))
    want = [sum(dense[i][j] * z[j] for j in range(cols)) % pm.R for i in range(rows)]
    assert _canon(out) == want
