"""liblightzk_amd — MI355X-native Groth16 prover for alt_bn128 (Python host side).

Thin Python mirror over the C ABI in include/hzk.h (libhzk.so, hand-written HIP
kernels for gfx950). It keeps the names and argument meaning of LibLightZK's
operator and protocol API for the proving path:

  multi_scalar_multiplication_g1/_g2   include/operator/msm.cuh:210-248
  number_theory_transformation         include/operator/poly.cuh:65-73
  batch_sparse_matrix_vector_multiplication  poly.cuh:5-30
  polynomial_A_times_B_minus_C         poly.cuh:75-86
  ntt_setup                            include/protocol/groth16-operator.cuh:6-32
  Groth16 (Setup / Prove)              include/protocol/groth16.cuh:18-223

Device memory is held in DeviceBuffer objects (the role of xpu::vector,
include/data-structure/xpu-vector.cuh:10-101). Every call runs on the GPU;
there is no CPU fallback, and a missing library raises HzkError.
"""
import ctypes
import os

import numpy as np

from ._lib import HzkError, load

__all__ = ["HzkError", "Context", "DeviceBuffer", "load"]

G1_AFFINE_BYTES = 64
G2_AFFINE_BYTES = 128
G1_JAC_BYTES = 96
G2_JAC_BYTES = 192
FIELD_BYTES = 32


class MsmDesc(ctypes.Structure):
    """hzk_msm_desc: one instance of a batched MSM."""
    _fields_ = [("scalars", ctypes.c_void_p), ("bases", ctypes.c_void_p), ("n", ctypes.c_size_t),
                ("dest", ctypes.c_void_p), ("win_siz", ctypes.c_uint32), ("levels", ctypes.c_uint32)]


class NttDesc(ctypes.Structure):
    """hzk_ntt_desc: one transform of a batch (kind 0 NTT, 1 iNTT, 2 coset NTT, 3 coset iNTT)."""
    _fields_ = [("coeffs", ctypes.c_void_p), ("deg", ctypes.c_size_t), ("kind", ctypes.c_int)]


class DeviceBuffer:
    """Owned device allocation (xpu::vector with mem_policy::device_only)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._check(load().hzk_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def __int__(self):
        return self.ptr

    def offset(self, nbytes):
        """Raw device address nbytes into this buffer."""
        return self.ptr + int(nbytes)

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes, "upload overflows buffer"
        self.ctx._check(load().hzk_memcpy_h2d(self.ctx.handle, self.ptr + offset, a.ctypes.data, a.nbytes))
        return self

    def download(self, dtype=np.uint64, count=None, offset=0):
        dt = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dt.itemsize
        out = np.empty(count, dtype=dt)
        assert offset + out.nbytes <= self.nbytes, "download overflows buffer"
        self.ctx._check(load().hzk_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr + offset, out.nbytes))
        return out

    def zero(self):
        self.ctx._check(load().hzk_memset(self.ctx.handle, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            load().hzk_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.free()
        except Exception:
            pass


class Context:
    """One device + one HIP stream + scratch arena (hzk_ctx)."""

    def __init__(self, device=0, stream=None):
        L = load()
        h = ctypes.c_void_p()
        rc = L.hzk_ctx_create(device, stream, ctypes.byref(h))
        if rc != 0:
            raise HzkError(f"hzk_ctx_create(device={device}) failed rc={rc}: no usable GPU")
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", None):
            load().hzk_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            msg = load().hzk_last_error(self.handle)
            raise HzkError(f"rc={rc}: {msg.decode() if msg else ''}")
        return rc

    # ---- buffers
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        return DeviceBuffer(self, max(a.nbytes, 1)).upload(a)

    def sync(self):
        self._check(load().hzk_sync(self.handle))

    def mem_info(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(load().hzk_mem_info(self.handle, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    # ---- timing
    def timing(self, enable=True):
        self._check(load().hzk_timing_enable(self.handle, 1 if enable else 0))

    def timing_reset(self):
        self._check(load().hzk_timing_reset(self.handle))

    def timing_get(self, name):
        ms, n = ctypes.c_double(), ctypes.c_long()
        self._check(load().hzk_timing_get(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def timing_report(self):
        need = load().hzk_timing_report(self.handle, None, 0)
        buf = ctypes.create_string_buffer(need + 1)
        load().hzk_timing_report(self.handle, buf, need + 1)
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n = line.split()
            out[name] = (float(ms), int(n))
        return out

    def roofline_probe(self, kind):
        """kind: 'mad' (v_mad_u64_u32/s), 'fqmul' (Fq products/s), 'add' (VALU adds/s), 'copy' (HBM bytes/s),
        'dbl_chain_g1' / 'dbl_chain_g2' (dependent point doublings/s on one lane of a single wave)."""
        v = ctypes.c_double()
        code = {"mad": 0, "fqmul": 1, "add": 2, "copy": 3, "dbl_chain_g1": 4, "dbl_chain_g2": 5}[kind]
        self._check(load().hzk_roofline_probe(self.handle, code, ctypes.byref(v)))
        return v.value

    def set_tuning(self, window_bits=0, segment=0):
        self._check(load().hzk_set_tuning(self.handle, window_bits, segment))

    def set_msm_regime(self, regime):
        """-1 by batch width (default), 0 single-MSM plans, 1 wide-batch plans."""
        self._check(load().hzk_set_msm_regime(self.handle, regime))

    def msm_tune_set(self, group, log_n, c, K=0, g=0):
        self._check(load().hzk_msm_tune_set(self.handle, group, log_n, c, K, g))

    def msm_tune_get(self, group, log_n):
        c, K, g = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_uint32()
        self._check(load().hzk_msm_tune_get(self.handle, group, log_n, ctypes.byref(c), ctypes.byref(K), ctypes.byref(g)))
        return c.value, K.value, g.value

    def msm_autotune(self, group, log_min, log_max, budget_ms=2000.0):
        self._check(load().hzk_msm_autotune(self.handle, group, log_min, log_max, float(budget_ms)))
        return {lg: self.msm_tune_get(group, lg) for lg in range(log_min, log_max + 1)}

    def msm_batch_tune_set(self, group, log_n, c, K=0, g=0):
        """Plan of a size class in the wide-batch (throughput) regime."""
        self._check(load().hzk_msm_batch_tune_set(self.handle, group, log_n, c, K, g))

    def msm_batch_tune_get(self, group, log_n):
        c, K, g = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_uint32()
        self._check(load().hzk_msm_batch_tune_get(self.handle, group, log_n, ctypes.byref(c), ctypes.byref(K), ctypes.byref(g)))
        return c.value, K.value, g.value

    def msm_batch_autotune(self, group, log_min, log_max, budget_ms=2000.0):
        self._check(load().hzk_msm_batch_autotune(self.handle, group, log_min, log_max, float(budget_ms)))
        return {lg: self.msm_batch_tune_get(group, lg) for lg in range(log_min, log_max + 1)}

    def inject_alloc_failure(self, after_n):
        self._check(load().hzk_inject_alloc_failure(self.handle, after_n))

    # ---- operators (device pointers in, device pointers out)
    def field_op(self, field, op, a, b, out, n):
        fields = {"fq": 0, "fr": 1, "fq2": 2}
        ops = {"mul": 0, "add": 1, "sub": 2, "neg": 3, "sqr": 4, "inv": 5, "from_mont": 6, "to_mont": 7}
        self._check(load().hzk_field_op(self.handle, fields[field], ops[op], int(a), int(b) if b is not None else None,
                                        int(out), n))

    def ntt(self, coeffs, deg, omega_powers=None, inverse=False, factor=None):
        self._check(load().hzk_ntt(self.handle, int(coeffs), deg, int(omega_powers) if omega_powers is not None else None,
                                   1 if inverse else 0, int(factor) if factor is not None else None))

    def coset_ntt(self, coeffs, deg, inverse=False):
        self._check(load().hzk_coset_ntt(self.handle, int(coeffs), deg, 1 if inverse else 0))

    def msm_batch(self, group, jobs, host_dest=False, graph=False, device_combine=False):
        """jobs: list of (scalars, bases, n, dest[, win_siz[, levels]]); device pointers (dest host if
        host_dest); levels != 0 marks bases as a level table built by bases_levels for win_siz.
        device_combine: the window combine runs on the device and the call returns without waiting."""
        arr = (MsmDesc * len(jobs))()
        for i, j in enumerate(jobs):
            arr[i] = MsmDesc(int(j[0]), int(j[1]), j[2], int(j[3]), j[4] if len(j) > 4 else 0, j[5] if len(j) > 5 else 0)
        flags = (1 if host_dest else 0) | (2 if graph else 0) | (4 if device_combine else 0)
        fn = load().hzk_msm_g1_batch if group == 1 else load().hzk_msm_g2_batch
        self._check(fn(self.handle, arr, len(jobs), flags))

    def ntt_batch(self, jobs, graph=False):
        """jobs: list of (device pointer, deg, kind)."""
        arr = (NttDesc * len(jobs))()
        for i, (p, deg, kind) in enumerate(jobs):
            arr[i] = NttDesc(int(p), deg, kind)
        self._check(load().hzk_ntt_batch(self.handle, arr, len(jobs), 2 if graph else 0))

    def msm_last_stats(self):
        a = (ctypes.c_uint32 * 5)()
        self._check(load().hzk_msm_last_stats(self.handle, a))
        return dict(zip(("entries", "chunks", "max_bucket", "c", "windows"), list(a)))

    def fixbase_g1(self, scalars, n, out_affine):
        self._check(load().hzk_fixbase_g1(self.handle, int(scalars), n, int(out_affine)))

    def fixbase_g2(self, scalars, n, out_affine):
        self._check(load().hzk_fixbase_g2(self.handle, int(scalars), n, int(out_affine)))

    def msm_g1(self, scalars, bases, n, dest, win_siz=0, layout=0):
        self._check(load().hzk_msm_g1(self.handle, int(scalars), int(bases), n, layout, win_siz, int(dest)))

    def msm_g1_noalloc(self, scalars, bases, n, dest):
        """hzk_msm_g1_noalloc: the G1 MSM on the library's static device scratch (no allocation)."""
        self._check(load().hzk_msm_g1_noalloc(self.handle, int(scalars), int(bases), n, int(dest)))

    def bases_levels(self, group, c, bases, n, out, levels=None):
        """Level table out[j*n + i] = 2^(c j) bases[i] (affine), j < levels (default ceil(255 / c))."""
        levels = levels or (255 + c - 1) // c
        self._check(load().hzk_bases_levels(self.handle, group, c, levels, int(bases), n, int(out)))

    def msm_g2(self, scalars, bases, n, dest, win_siz=0, layout=0):
        self._check(load().hzk_msm_g2(self.handle, int(scalars), int(bases), n, layout, win_siz, int(dest)))


class Csr(ctypes.Structure):
    """hzk_csr: the reference's spmat_descript (xpu-spmat.cuh:20-25) over device arrays."""
    _fields_ = [("val", ctypes.c_void_p), ("col_idx", ctypes.c_void_p), ("row_ptr", ctypes.c_void_p),
                ("row", ctypes.c_size_t), ("col", ctypes.c_size_t), ("cnt", ctypes.c_size_t)]


def batch_sparse_matrix_vector_multiplication(ctx, A, B, C, z, Az, Bz, Cz):
    """Reference operator poly.cuh:5-30: Az, Bz, Cz <- A z, B z, C z (Csr descriptors, device vectors)."""
    ctx._check(load().hzk_spmv3(ctx.handle, ctypes.byref(A), ctypes.byref(B), ctypes.byref(C), int(z), int(Az),
                                int(Bz), int(Cz)))


def polynomial_A_times_B_minus_C(ctx, deg, A, B, C):
    """Reference operator poly.cuh:75-86: A <- A o B - C."""
    ctx._check(load().hzk_pointwise_abc(ctx.handle, deg, int(A), int(B), int(C)))


class Groth16:
    """Reference protocol class (include/protocol/groth16.cuh:18-223) on the GPU.

    Groth16(ctx, seed, log_m) builds the synthetic circuit and runs Setup (keys
    resident in HBM); prove(z, rs) returns the 256-byte proof A | B | C."""

    def __init__(self, ctx, seed, log_m=None, num_public=8, matmul_p=None, path=None, _handle=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if path is not None:      # proving key + R1CS from a file (hzk_g16_load)
            ctx._check(load().hzk_g16_load(ctx.handle, os.fsencode(path), ctypes.byref(h)))
        elif matmul_p is not None:  # the reference's MatMul statement (input_scale p)
            ctx._check(load().hzk_g16_setup_matmul(ctx.handle, seed, matmul_p, ctypes.byref(h)))
        else:
            ctx._check(load().hzk_g16_setup(ctx.handle, seed, log_m, num_public, ctypes.byref(h)))
        self.h = h.value
        info = (ctypes.c_uint64 * 8)()
        load().hzk_g16_info(self.h, info)
        keys = ["M", "num_vars", "num_public", "num_gates", "nnzA", "nnzB", "nnzC", "key_bytes"]
        self.info = {k: int(v) for k, v in zip(keys, info)}

    def close(self):
        if getattr(self, "h", None):
            load().hzk_g16_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def witness(self, wseed):
        z = np.empty(4 * self.info["num_vars"], dtype=np.uint64)
        load().hzk_g16_witness(self.h, wseed, z.ctypes.data)
        return z

    @staticmethod
    def rs(wseed):
        out = np.empty(8, dtype=np.uint64)
        load().hzk_g16_rs(wseed, out.ctypes.data)
        return out

    def prove(self, z, rs):
        z = np.ascontiguousarray(z, dtype=np.uint64)
        rs = np.ascontiguousarray(rs, dtype=np.uint64)
        proof = np.empty(32, dtype=np.uint64)
        self.ctx._check(load().hzk_g16_prove(self.ctx.handle, self.h, z.ctypes.data, rs.ctypes.data, proof.ctypes.data))
        return proof

    def set_levels(self, levels):
        """Prove over level tables with `levels` levels per bucket set (<= 1: plain key arrays)."""
        self.ctx._check(load().hzk_g16_set_levels(self.ctx.handle, self.h, levels))

    @classmethod
    def matmul(cls, ctx, seed, p):
        """The reference's MatMul circuit for p x p matrices (input_scale p = 39 / 63 / 100)."""
        return cls(ctx, seed, matmul_p=p)

    @classmethod
    def load_file(cls, ctx, path):
        """Proving key and R1CS written by save()."""
        return cls(ctx, 0, path=path)

    def save(self, path):
        self.ctx._check(load().hzk_g16_save(self.ctx.handle, self.h, os.fsencode(path)))

    def key_array(self, which):
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        self.ctx._check(load().hzk_g16_key_array(self.h, which, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value


class ProofServer:
    """Continuous-batching Groth16 proof server (hzk_server_*).

    add_circuit() before start(); submit() may be called from any thread and
    returns a request id; wait() returns (status, proof, latency_s) with status
    1 done, 0 still running, < 0 failed (an HZK_ERR_* code)."""

    def __init__(self, device=0, mem_budget=0, max_batch=64, mode="continuous", pipelines=3):
        h = ctypes.c_void_p()
        m = {"continuous": 0, "sequential": 1}[mode]
        rc = load().hzk_server_create(device, mem_budget, max_batch, m, ctypes.byref(h))
        if rc != 0:
            raise HzkError(f"hzk_server_create rc={rc}")
        self.h = h.value
        self.circuits = []
        self._check(load().hzk_server_set_pipelines(self.h, pipelines))

    def _check(self, rc):
        if rc != 0:
            msg = load().hzk_server_last_error(self.h)
            raise HzkError(f"rc={rc}: {msg.decode() if msg else ''}")

    def add_circuit(self, seed, log_m, num_public=8):
        cid = ctypes.c_int()
        self._check(load().hzk_server_add_circuit(self.h, seed, log_m, num_public, ctypes.byref(cid)))
        g = ctypes.c_void_p()
        self._check(load().hzk_server_circuit(self.h, cid.value, ctypes.byref(g)))
        info = (ctypes.c_uint64 * 8)()
        load().hzk_g16_info(g.value, info)
        keys = ["M", "num_vars", "num_public", "num_gates", "nnzA", "nnzB", "nnzC", "key_bytes"]
        self.circuits.append((g.value, {k: int(v) for k, v in zip(keys, info)}))
        return cid.value

    def add_circuit_file(self, path):
        cid = ctypes.c_int()
        self._check(load().hzk_server_add_circuit_file(self.h, os.fsencode(path), ctypes.byref(cid)))
        g = ctypes.c_void_p()
        self._check(load().hzk_server_circuit(self.h, cid.value, ctypes.byref(g)))
        info = (ctypes.c_uint64 * 8)()
        load().hzk_g16_info(g.value, info)
        keys = ["M", "num_vars", "num_public", "num_gates", "nnzA", "nnzB", "nnzC", "key_bytes"]
        self.circuits.append((g.value, {k: int(v) for k, v in zip(keys, info)}))
        return cid.value

    def set_levels(self, cid, levels):
        """Level tables for a circuit's key (before start); proofs are unchanged."""
        self._check(load().hzk_server_set_levels(self.h, cid, levels))

    def witness(self, cid, wseed):
        g, info = self.circuits[cid]
        z = np.empty(4 * info["num_vars"], dtype=np.uint64)
        load().hzk_g16_witness(g, wseed, z.ctypes.data)
        return z

    def footprint(self, cid):
        n = ctypes.c_size_t()
        self._check(load().hzk_server_footprint(self.h, cid, ctypes.byref(n)))
        return n.value

    def start(self):
        self._check(load().hzk_server_start(self.h))

    def set_mode(self, mode):
        self._check(load().hzk_server_set_mode(self.h, {"continuous": 0, "sequential": 1}[mode]))

    def submit(self, cid, z, rs):
        z = np.ascontiguousarray(z, dtype=np.uint64)
        rs = np.ascontiguousarray(rs, dtype=np.uint64)
        rid = ctypes.c_uint64()
        self._check(load().hzk_server_submit(self.h, cid, z.ctypes.data, rs.ctypes.data, ctypes.byref(rid)))
        return rid.value

    def wait(self, rid, timeout=600.0):
        proof = np.empty(32, dtype=np.uint64)
        lat = ctypes.c_double()
        st = load().hzk_server_wait(self.h, rid, float(timeout), proof.ctypes.data, ctypes.byref(lat))
        return st, (proof if st == 1 else None), lat.value

    def stats(self):
        out = (ctypes.c_double * 7)()
        self._check(load().hzk_server_stats(self.h, out))
        keys = ["steps", "completed", "failed", "max_inflight", "mean_batch", "budget", "in_use"]
        return {k: float(v) for k, v in zip(keys, out)}

    def inject_failure(self, after_n):
        self._check(load().hzk_server_inject_failure(self.h, after_n))

    def last_error(self):
        msg = load().hzk_server_last_error(self.h)
        return msg.decode() if msg else ""

    def close(self):
        if getattr(self, "h", None):
            load().hzk_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_field(seed, stream, n, kind="fr"):
    """Seeded synthetic field elements (host numpy uint64, 4 limbs each)."""
    out = np.empty(4 * n, dtype=np.uint64)
    code = {"fr": 0, "fr_sparse": 1, "fq": 2}[kind]
    rc = load().hzk_gen_field(seed, stream, n, code, out.ctypes.data)
    if rc != 0:
        raise HzkError(f"hzk_gen_field rc={rc}")
    return out


def ntt_setup(M):
    """The reference's ntt_setup table (include/protocol/groth16-operator.cuh:6-32)
    for an M-point domain: 6M + 2 Fr elements on the host (numpy uint64, 4 limbs each)."""
    if M < 1 or M & (M - 1) or M > 1 << 26:
        raise HzkError(f"ntt_setup: M={M} must be a power of two <= 2^26")
    out = np.empty(4 * (6 * M + 2), dtype=np.uint64)
    rc = load().hzk_ntt_setup(M, out.ctypes.data)
    if rc != 0:
        raise HzkError(f"hzk_ntt_setup(M={M}) rc={rc}: M must be a power of two <= 2^26")
    return out


def number_theory_transformation(ctx, deg, log_deg, coeffs, omega_powers, factor=None):
    """Reference operator poly.cuh:65-73: coeffs <- factor * DFT(coeffs) with w = omega_powers[1]."""
    if deg != 1 << log_deg:
        raise HzkError("deg must equal 2^log_deg")
    ctx.ntt(coeffs, deg, omega_powers, False, factor)


def multi_scalar_multiplication_g1(ctx, length, scalars, points, dest, win_siz=0, win_cnt=None, bucket_cnt=None,
                                   level_stride=None):
    """Reference operator msm.cuh:210-214 over affine bases: dest <- sum scalars[i] * points[i].

    win_cnt, bucket_cnt and level_stride are accepted for signature parity and
    derived from win_siz (the reference hard-wires c = 12 geometry; here any c
    in 4..20, or 0 for the autotuned choice)."""
    ctx.msm_g1(scalars, points, length, dest, win_siz)


def multi_scalar_multiplication_g2(ctx, length, scalars, points, dest, win_siz=0, win_cnt=None, bucket_cnt=None,
                                   level_stride=None):
    """Reference operator msm.cuh:230-234 for G2 (affine Fq2 bases, 192-byte Jacobian dest)."""
    ctx.msm_g2(scalars, points, length, dest, win_siz)
