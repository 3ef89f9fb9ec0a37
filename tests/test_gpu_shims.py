"""The reference-shaped C++ API (include/operator/msm.hpp, poly.hpp,
include/protocol/groth16.hpp) driven by tests/cpp/shim_check the way the
reference's own tests call the operators, on seeded inputs written here.

Checks: fixed-base generation + pre_comp + MSM in the reference's record
layout give the oracle's closed form (sum s_i k_i) G for G1 and G2; the
pre_comp level records are 2^(gap j) P_0 with Z = 1; the NTT with the
caller's ntt_setup table and factor equals the oracle's restatement of the
reference operator; SpMV equals the oracle's CSR product and A o B - C
vanishes on a satisfying witness; the protocol class (input_scale p = 39,
the reference's MatMul statement) proves a witness equal to the oracle's and
a proof equal to the C ABI prover's and to the oracle's trapdoor closed form;
a bad call raises lightzk::error with the library's message."""
import os
import subprocess

import numpy as np
import pytest

import pymodel as pm

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "shim_check")


def _write(d, name, arr):
    np.ascontiguousarray(arr).tofile(os.path.join(d, name))


def _read(d, name, dtype=np.uint64):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


@pytest.fixture(scope="module")
def run(orc, tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/shim_check is not built (run __graft_entry__.build())")
    d = str(tmp_path_factory.mktemp("shim"))
    n, gap, level, deg = 3000, 12, 3, 1 << 12
    s, k = orc.gen_fr(71, 0, n), orc.gen_fr(71, 1, n)
    x = orc.gen_fr(71, 2, deg)
    tab = orc.ntt_setup(deg)
    seed, log_m, npub = 7, 8, 8
    sizes = orc.circuit_sizes(seed, log_m, npub)
    M, nv = sizes["M"], sizes["num_vars"]
    zz = orc.circuit_witness(seed, log_m, npub, 77)
    nnz = []
    for which, nm in enumerate("abc"):
        rp, col, val = orc.circuit_csr(seed, log_m, npub, which)
        _write(d, f"csr_{nm}_val.bin", val)
        _write(d, f"csr_{nm}_col.bin", col)
        _write(d, f"csr_{nm}_rp.bin", rp)
        nnz.append(col.size)
    _write(d, "scalars.bin", s)
    _write(d, "ks.bin", k)
    _write(d, "ntt_in.bin", x)
    _write(d, "ntt_tab.bin", tab[: 4 * deg])
    _write(d, "ntt_factor.bin", tab[8 * deg: 8 * deg + 4])   # 1/deg: the inverse table's scale
    _write(d, "z.bin", zz)
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write(f"{n} {gap} {level} {deg} {M} {nv} {nnz[0]} {nnz[1]} {nnz[2]}\n")
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=600)
    status = open(os.path.join(d, "status.txt")).read().split("\n")
    assert r.returncode == 0 and status[0] == "ok", (r.returncode, r.stderr, status)
    return dict(d=d, n=n, gap=gap, level=level, deg=deg, s=s, k=k, x=x, tab=tab, seed=seed, log_m=log_m, npub=npub,
                M=M, nv=nv, zz=zz, status=status)


def test_msm_reference_layout(orc, run):
    d, s, k = run["d"], run["s"], run["k"]
    assert np.array_equal(_read(d, "msm_g1.bin"), orc.msm_closed_form(s, k, 1))
    assert np.array_equal(_read(d, "msm_g2.bin"), orc.msm_closed_form(s, k, 2))


def test_pre_comp_levels(orc, run):
    d, gap, level, k = run["d"], run["gap"], run["level"], run["k"]
    one = pm.int_to_limbs(pm.MONT % pm.P)
    lv1 = _read(d, "levels_g1.bin").reshape(level, 12)
    lv2 = _read(d, "levels_g2.bin").reshape(level, 24)
    k0 = pm.from_mont(pm.limbs_to_int(k[:4]), pm.R)
    for j in range(level):
        kj = np.array(pm.int_to_limbs(pm.to_mont(k0 * pow(2, gap * j, pm.R) % pm.R, pm.R)), dtype=np.uint64)
        aff1 = orc.g1_mul_gen(kj)                      # 2^(gap j) P_0, affine x | y
        assert np.array_equal(lv1[j][:8], aff1[:8]), j
        assert np.array_equal(lv1[j][8:], one), j     # Z = 1
        aff2 = orc.g2_mul_gen(kj)                      # x.c0 | x.c1 | y.c0 | y.c1
        rec0, rec1 = lv2[j][:12], lv2[j][12:]          # lane records [X.c0|Y.c0|Z.c0] [X.c1|Y.c1|Z.c1]
        assert np.array_equal(np.concatenate([rec0[:4], rec1[:4], rec0[4:8], rec1[4:8]]), aff2[:16]), j
        assert np.array_equal(rec0[8:], one) and not rec1[8:].any(), j


def test_ntt_and_spmv(orc, run):
    d, deg, x, tab = run["d"], run["deg"], run["x"], run["tab"]
    want = orc.ntt(x, tab[: 4 * deg], tab[8 * deg: 8 * deg + 4].copy())
    assert np.array_equal(_read(d, "ntt_out.bin"), want)
    M = run["M"]
    got = _read(d, "spmv.bin").reshape(3, 4 * M)
    for which in range(3):
        rp, col, val = orc.circuit_csr(run["seed"], run["log_m"], run["npub"], which)
        assert np.array_equal(got[which], orc.spmv(M, col.size, val, col, rp, run["zz"])), which
    assert not _read(d, "abc.bin").any()


def test_protocol_class_and_errors(ctx, orc, run):
    import liblightzk_amd as z
    g = z.Groth16.matmul(ctx, 2, 39)      # input_scale::MatMul_p_39
    zz, rs = g.witness(1), g.rs(1)
    want = g.prove(zz, rs)
    g.close()
    got = _read(run["d"], "g16_proof.bin")
    assert np.array_equal(got, want)
    o = orc.Groth16(2, matmul_p=39)       # the reference's statement at its own size
    assert np.array_equal(zz, o.witness(1))
    assert np.array_equal(got, o.prove_trapdoor(zz, rs))
    assert "not a power of two" in run["status"][1] or "deg" in run["status"][1], run["status"]
