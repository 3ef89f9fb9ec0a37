"""ORACLE — test infrastructure only.

ctypes binding of oracle/libhzk_oracle.so, the CPU restatement of the
reference's host arithmetic (libff subset) and of the operator semantics of
include/operator/{msm,poly}.cuh. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker.

All buffers are numpy uint64 arrays in the byte layouts documented in
orc_core.cpp (32-byte Montgomery field elements, 64/128-byte affine points,
96/192-byte normalised Jacobian points).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U64P = ctypes.POINTER(ctypes.c_uint64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libhzk_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        sz, u64, i32, vp = ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        sigs = {
            "orc_fq_op": (i32, [i32, U64P, U64P, U64P, sz]),
            "orc_fr_op": (i32, [i32, U64P, U64P, U64P, sz]),
            "orc_fq2_op": (i32, [i32, U64P, U64P, U64P, sz]),
            "orc_g1_op": (i32, [i32, U64P, U64P, U64P, sz]),
            "orc_g2_op": (i32, [i32, U64P, U64P, U64P, sz]),
            "orc_g1_on_curve": (i32, [U64P]),
            "orc_g2_on_curve": (i32, [U64P]),
            "orc_g1_generator": (None, [U64P]),
            "orc_g2_generator": (None, [U64P]),
            "orc_g1_mul_gen": (i32, [U64P, U64P, sz]),
            "orc_g2_mul_gen": (i32, [U64P, U64P, sz]),
            "orc_g1_scalar_mul": (i32, [U64P, U64P, U64P]),
            "orc_g2_scalar_mul": (i32, [U64P, U64P, U64P]),
            "orc_msm_g1": (i32, [U64P, U64P, sz, i32, U64P]),
            "orc_msm_g2": (i32, [U64P, U64P, sz, i32, U64P]),
            "orc_msm_closed_form": (i32, [U64P, U64P, sz, i32, U64P]),
            "orc_gen_fr": (i32, [u64, u64, sz, i32, U64P]),
            "orc_gen_fq": (i32, [u64, u64, sz, U64P]),
            "orc_rng_u64": (u64, [u64, u64, sz, U64P]),
            "orc_ntt": (i32, [sz, sz, U64P, U64P, U64P]),
            "orc_fr_root_of_unity": (i32, [i32, i32, U64P]),
            "orc_ntt_setup": (i32, [sz, U64P]),
            "orc_spmv": (i32, [sz, sz, U64P, U64P, U64P, U64P, U64P]),
            "orc_num_threads": (i32, []),
            "orc_pairing": (i32, [U64P, U64P, U64P]),
            "orc_pairing_product_is_one": (i32, [U64P, U64P, sz]),
            "orc_twist_frobenius_consts": (None, [U64P]),
            "orc_circuit_sizes": (i32, [u64, i32, i32, U64P]),
            "orc_circuit_csr": (i32, [u64, i32, i32, i32, U64P, U64P, U64P]),
            "orc_circuit_witness": (i32, [u64, i32, i32, u64, U64P]),
            "orc_r1cs_satisfied": (i32, [u64, i32, i32, U64P]),
            "orc_g16_new": (vp, [u64, i32, i32]),
            "orc_circuit_sizes_k": (i32, [i32, u64, i32, i32, U64P]),
            "orc_circuit_csr_k": (i32, [i32, u64, i32, i32, i32, U64P, U64P, U64P]),
            "orc_circuit_witness_k": (i32, [i32, u64, i32, i32, u64, U64P]),
            "orc_r1cs_satisfied_k": (i32, [i32, u64, i32, i32, U64P]),
            "orc_g16_new_k": (vp, [i32, u64, i32, i32]),
            "orc_g16_free": (None, [vp]),
            "orc_g16_key_scalars": (sz, [vp, i32, U64P]),
            "orc_g16_key_points": (sz, [vp, i32, U64P]),
            "orc_g16_h": (i32, [vp, U64P, U64P]),
            "orc_g16_rs": (None, [u64, U64P]),
            "orc_g16_prove_trapdoor": (i32, [vp, U64P, U64P, U64P]),
            "orc_g16_prove": (i32, [vp, U64P, U64P, U64P]),
            "orc_g16_verify": (i32, [vp, U64P, U64P]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(U64P)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle call failed rc={rc}")


# ---------------------------------------------------------------- fields
def field_op(kind, op, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    width = 8 if kind == "fq2" else 4
    n = a.size // width
    out = np.zeros_like(a)
    fn = {"fq": lib().orc_fq_op, "fr": lib().orc_fr_op, "fq2": lib().orc_fq2_op}[kind]
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.uint64)
    _chk(fn(op, ptr(a), ptr(bb), ptr(out), n))
    return out


def gen_fr(seed, stream, n, sparse=False):
    out = np.zeros(4 * n, dtype=np.uint64)
    _chk(lib().orc_gen_fr(seed, stream, n, 1 if sparse else 0, ptr(out)))
    return out


def gen_fq(seed, stream, n):
    out = np.zeros(4 * n, dtype=np.uint64)
    _chk(lib().orc_gen_fq(seed, stream, n, ptr(out)))
    return out


def rng_u64(seed, stream, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_rng_u64(seed, stream, n, ptr(out))
    return out


# ---------------------------------------------------------------- groups
def g1_mul_gen(k):
    n = k.size // 4
    out = np.zeros(8 * n, dtype=np.uint64)
    _chk(lib().orc_g1_mul_gen(ptr(k), ptr(out), n))
    return out


def g2_mul_gen(k):
    n = k.size // 4
    out = np.zeros(16 * n, dtype=np.uint64)
    _chk(lib().orc_g2_mul_gen(ptr(k), ptr(out), n))
    return out


def msm_g1(scalars, bases_aff, method=1):
    n = scalars.size // 4
    out = np.zeros(12, dtype=np.uint64)
    _chk(lib().orc_msm_g1(ptr(scalars), ptr(bases_aff), n, method, ptr(out)))
    return out


def msm_g2(scalars, bases_aff, method=1):
    n = scalars.size // 4
    out = np.zeros(24, dtype=np.uint64)
    _chk(lib().orc_msm_g2(ptr(scalars), ptr(bases_aff), n, method, ptr(out)))
    return out


def msm_closed_form(scalars, ks, group=1):
    n = scalars.size // 4
    out = np.zeros(12 if group == 1 else 24, dtype=np.uint64)
    _chk(lib().orc_msm_closed_form(ptr(scalars), ptr(ks), n, group, ptr(out)))
    return out


def g1_generator():
    out = np.zeros(12, dtype=np.uint64)
    lib().orc_g1_generator(ptr(out))
    return out


def g2_generator():
    out = np.zeros(24, dtype=np.uint64)
    lib().orc_g2_generator(ptr(out))
    return out


# ---------------------------------------------------------------- NTT / SpMV
def ntt(coeffs, omega_powers, factor=None):
    a = np.ascontiguousarray(coeffs, dtype=np.uint64).copy()
    deg = a.size // 4
    log_deg = deg.bit_length() - 1
    _chk(lib().orc_ntt(deg, log_deg, ptr(a), ptr(omega_powers), ptr(factor)))
    return a


def root_of_unity(k, inverse=False):
    out = np.zeros(4, dtype=np.uint64)
    _chk(lib().orc_fr_root_of_unity(k, 1 if inverse else 0, ptr(out)))
    return out


def ntt_setup(M):
    out = np.zeros(4 * (6 * M + 2), dtype=np.uint64)
    _chk(lib().orc_ntt_setup(M, ptr(out)))
    return out


def spmv(row, cnt, val, col_idx, row_ptr, z):
    out = np.zeros(4 * row, dtype=np.uint64)
    _chk(lib().orc_spmv(row, cnt, ptr(val), ptr(col_idx), ptr(row_ptr), ptr(z), ptr(out)))
    return out


# ---------------------------------------------------------------- pairing
def pairing(p_aff, q_aff):
    out = np.zeros(48, dtype=np.uint64)
    _chk(lib().orc_pairing(ptr(p_aff), ptr(q_aff), ptr(out)))
    return out


def pairing_product_is_one(ps, qs):
    return bool(lib().orc_pairing_product_is_one(ptr(ps), ptr(qs), ps.size // 8))


# ---------------------------------------------------------------- Groth16
def circuit_sizes(seed, log_m, num_public):
    out = np.zeros(7, dtype=np.uint64)
    _chk(lib().orc_circuit_sizes(seed, log_m, num_public, ptr(out)))
    keys = ["M", "num_vars", "num_public", "num_gates", "nnzA", "nnzB", "nnzC"]
    return {k: int(v) for k, v in zip(keys, out)}


def circuit_csr(seed, log_m, num_public, which):
    s = circuit_sizes(seed, log_m, num_public)
    nnz = s["nnz" + "ABC"[which]]
    row_ptr = np.zeros(s["M"], dtype=np.uint64)
    col = np.zeros(nnz, dtype=np.uint64)
    val = np.zeros(4 * nnz, dtype=np.uint64)
    _chk(lib().orc_circuit_csr(seed, log_m, num_public, which, ptr(row_ptr), ptr(col), ptr(val)))
    return row_ptr, col, val


def circuit_witness(seed, log_m, num_public, wseed):
    s = circuit_sizes(seed, log_m, num_public)
    z = np.zeros(4 * s["num_vars"], dtype=np.uint64)
    _chk(lib().orc_circuit_witness(seed, log_m, num_public, wseed, ptr(z)))
    return z


def r1cs_satisfied(seed, log_m, num_public, z):
    return bool(lib().orc_r1cs_satisfied(seed, log_m, num_public, ptr(z)))


def matmul_sizes(p):
    """Sizes of the reference's MatMul circuit for p x p matrices."""
    out = np.zeros(7, dtype=np.uint64)
    _chk(lib().orc_circuit_sizes_k(1, 0, p, 0, ptr(out)))
    keys = ["M", "num_vars", "num_public", "num_gates", "nnzA", "nnzB", "nnzC"]
    return {k: int(v) for k, v in zip(keys, out)}


def matmul_csr(p, which):
    s = matmul_sizes(p)
    nnz = s["nnz" + "ABC"[which]]
    row_ptr = np.zeros(s["M"], dtype=np.uint64)
    col = np.zeros(nnz, dtype=np.uint64)
    val = np.zeros(4 * nnz, dtype=np.uint64)
    _chk(lib().orc_circuit_csr_k(1, 0, p, 0, which, ptr(row_ptr), ptr(col), ptr(val)))
    return row_ptr, col, val


def matmul_witness(p, wseed):
    s = matmul_sizes(p)
    z = np.zeros(4 * s["num_vars"], dtype=np.uint64)
    _chk(lib().orc_circuit_witness_k(1, 0, p, 0, wseed, ptr(z)))
    return z


def matmul_satisfied(p, z):
    return bool(lib().orc_r1cs_satisfied_k(1, 0, p, 0, ptr(z)))


class Groth16:
    """Oracle Groth16 instance (circuit + key scalars; points on demand).
    matmul_p selects the reference's MatMul circuit instead of the
    multiplication-gate DAG."""

    def __init__(self, seed, log_m=None, num_public=8, matmul_p=None):
        self.seed, self.matmul_p = seed, matmul_p
        if matmul_p is not None:
            self.sizes = matmul_sizes(matmul_p)
            self.log_m = self.sizes["M"].bit_length() - 1
            self.num_public = self.sizes["num_public"]
            self.h = lib().orc_g16_new_k(1, seed, matmul_p, 0)
        else:
            self.log_m, self.num_public = log_m, num_public
            self.sizes = circuit_sizes(seed, log_m, num_public)
            self.h = lib().orc_g16_new(seed, log_m, num_public)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_g16_free(self.h)
            self.h = None

    def key_scalars(self, which):
        n = lib().orc_g16_key_scalars(self.h, which, None)
        out = np.zeros(4 * n, dtype=np.uint64)
        lib().orc_g16_key_scalars(self.h, which, ptr(out))
        return out

    def key_points(self, which):
        n = lib().orc_g16_key_points(self.h, which, None)
        width = 16 if which in (2,) else 8
        if which == 6:
            width = 8
        if which == 7:
            width = 16
        out = np.zeros(width * n, dtype=np.uint64)
        lib().orc_g16_key_points(self.h, which, ptr(out))
        return out

    def witness(self, wseed):
        if self.matmul_p is not None:
            return matmul_witness(self.matmul_p, wseed)
        return circuit_witness(self.seed, self.log_m, self.num_public, wseed)

    def h_coeffs(self, z):
        out = np.zeros(4 * self.sizes["M"], dtype=np.uint64)
        _chk(lib().orc_g16_h(self.h, ptr(z), ptr(out)))
        return out

    @staticmethod
    def rs(wseed):
        out = np.zeros(8, dtype=np.uint64)
        lib().orc_g16_rs(wseed, ptr(out))
        return out

    def prove_trapdoor(self, z, rs):
        out = np.zeros(32, dtype=np.uint64)
        _chk(lib().orc_g16_prove_trapdoor(self.h, ptr(z), ptr(rs), ptr(out)))
        return out

    def prove(self, z, rs):
        out = np.zeros(32, dtype=np.uint64)
        _chk(lib().orc_g16_prove(self.h, ptr(z), ptr(rs), ptr(out)))
        return out

    def verify(self, public_z, proof):
        pz = np.ascontiguousarray(public_z, dtype=np.uint64)
        return bool(lib().orc_g16_verify(self.h, ptr(pz), ptr(proof)))
