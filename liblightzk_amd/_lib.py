"""Loader for the in-tree C ABI library liblightzk_amd/libhzk.so (include/hzk.h).

The library is required: there is no CPU fallback anywhere in this package,
so a missing or unloadable library raises immediately.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhzk.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "hzk.h")

_lib = None


class HzkError(RuntimeError):
    pass


def header_functions(path=HEADER):
    """Names of every function the C ABI header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hzk_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HzkError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C liblightzk_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, u32, i64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32, ctypes.c_long
    dbl = ctypes.c_double
    sigs = {
        "hzk_ctx_create": (i32, [i32, vp, ctypes.POINTER(vp)]),
        "hzk_ctx_destroy": (None, [vp]),
        "hzk_last_error": (ctypes.c_char_p, [vp]),
        "hzk_sync": (i32, [vp]),
        "hzk_ctx_stream": (vp, [vp]),
        "hzk_version": (ctypes.c_char_p, []),
        "hzk_device_alloc": (i32, [vp, sz, ctypes.POINTER(vp)]),
        "hzk_device_free": (i32, [vp, vp]),
        "hzk_memcpy_h2d": (i32, [vp, vp, vp, sz]),
        "hzk_memcpy_d2h": (i32, [vp, vp, vp, sz]),
        "hzk_memset": (i32, [vp, vp, i32, sz]),
        "hzk_mem_info": (i32, [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]),
        "hzk_field_op": (i32, [vp, i32, i32, vp, vp, vp, sz]),
        "hzk_msm_g1": (i32, [vp, vp, vp, sz, u32, sz, vp]),
        "hzk_msm_g2": (i32, [vp, vp, vp, sz, u32, sz, vp]),
        "hzk_msm_g1_noalloc": (i32, [vp, vp, vp, sz, vp]),
        "hzk_ntt": (i32, [vp, vp, sz, vp, i32, vp]),
        "hzk_coset_ntt": (i32, [vp, vp, sz, i32]),
        "hzk_ntt_setup": (i32, [sz, vp]),
        "hzk_ntt_batch": (i32, [vp, vp, sz, u32]),
        "hzk_msm_g1_batch": (i32, [vp, vp, sz, u32]),
        "hzk_msm_g2_batch": (i32, [vp, vp, sz, u32]),
        "hzk_msm_last_stats": (i32, [vp, ctypes.POINTER(ctypes.c_uint32)]),
        "hzk_fixbase_g1": (i32, [vp, vp, sz, vp]),
        "hzk_fixbase_g2": (i32, [vp, vp, sz, vp]),
        "hzk_bases_convert": (i32, [vp, i32, u32, u32, vp, vp, sz]),
        "hzk_precomp": (i32, [vp, i32, vp, sz, sz, sz]),
        "hzk_bases_levels": (i32, [vp, i32, i32, i32, vp, sz, vp]),
        "hzk_device_sync": (i32, [vp]),
        "hzk_gen_field": (i32, [ctypes.c_uint64, ctypes.c_uint64, sz, i32, vp]),
        "hzk_spmv3": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "hzk_pointwise_abc": (i32, [vp, sz, vp, vp, vp]),
        "hzk_g16_setup": (i32, [vp, ctypes.c_uint64, i32, i32, ctypes.POINTER(vp)]),
        "hzk_g16_free": (None, [vp]),
        "hzk_g16_info": (i32, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "hzk_g16_witness": (i32, [vp, ctypes.c_uint64, vp]),
        "hzk_g16_rs": (i32, [ctypes.c_uint64, vp]),
        "hzk_g16_prove": (i32, [vp, vp, vp, vp, vp]),
        "hzk_g16_key_array": (i32, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
        "hzk_g16_set_levels": (i32, [vp, vp, i32]),
        "hzk_g16_setup_matmul": (i32, [vp, ctypes.c_uint64, i32, ctypes.POINTER(vp)]),
        "hzk_g16_save": (i32, [vp, vp, ctypes.c_char_p]),
        "hzk_g16_load": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "hzk_server_add_circuit_file": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(i32)]),
        "hzk_server_set_levels": (i32, [vp, i32, i32]),
        "hzk_server_create": (i32, [i32, sz, i32, i32, ctypes.POINTER(vp)]),
        "hzk_server_destroy": (None, [vp]),
        "hzk_server_set_pipelines": (i32, [vp, i32]),
        "hzk_server_last_error": (ctypes.c_char_p, [vp]),
        "hzk_server_add_circuit": (i32, [vp, ctypes.c_uint64, i32, i32, ctypes.POINTER(i32)]),
        "hzk_server_circuit": (i32, [vp, i32, ctypes.POINTER(vp)]),
        "hzk_server_footprint": (i32, [vp, i32, ctypes.POINTER(sz)]),
        "hzk_server_start": (i32, [vp]),
        "hzk_server_set_mode": (i32, [vp, i32]),
        "hzk_server_submit": (i32, [vp, i32, vp, vp, ctypes.POINTER(ctypes.c_uint64)]),
        "hzk_server_wait": (i32, [vp, ctypes.c_uint64, dbl, vp, ctypes.POINTER(dbl)]),
        "hzk_server_stats": (i32, [vp, ctypes.POINTER(dbl)]),
        "hzk_server_inject_failure": (i32, [vp, i64]),
        "hzk_timing_enable": (i32, [vp, i32]),
        "hzk_timing_reset": (i32, [vp]),
        "hzk_timing_get": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(dbl), ctypes.POINTER(i64)]),
        "hzk_timing_report": (sz, [vp, ctypes.c_char_p, sz]),
        "hzk_set_tuning": (i32, [vp, i32, i32]),
        "hzk_set_msm_regime": (i32, [vp, i32]),
        "hzk_msm_tune_set": (i32, [vp, i32, i32, i32, u32, u32]),
        "hzk_msm_tune_get": (i32, [vp, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "hzk_msm_autotune": (i32, [vp, i32, i32, i32, dbl]),
        "hzk_msm_batch_tune_set": (i32, [vp, i32, i32, i32, u32, u32]),
        "hzk_msm_batch_tune_get": (i32, [vp, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "hzk_msm_batch_autotune": (i32, [vp, i32, i32, i32, dbl]),
        "hzk_roofline_probe": (i32, [vp, i32, ctypes.POINTER(dbl)]),
        "hzk_inject_alloc_failure": (i32, [vp, i64]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L
