"""The reference's own test programs compiled against include/ (tests/cpp/ref_calls.cpp).

ref_calls keeps the call lines of test/msm-check.cu:17-41,
test/msm-g2-check.cu:24-51 and test/groth16.cu:5-8 (explicit template
arguments, xpu::vector buffers, libff host types, Groth16_BN254 on
input_scale::MatMul_p_100), plus ntt_setup and the MatMul xpu::spmat. The
reference checks its MSM against libff on the CPU (msm-check.cu:43-57); this
test checks every output against the oracle instead: the level-0 records of
the fixed-base table against k_i G, the MSM results against the closed form
(sum s_i k_i) G at the reference's n = 2^20, the ntt_setup table and the
MatMul CSR entry for entry, and the MatMul p = 100 proof against the trapdoor
closed form and the pairing verifier."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "ref_calls")
LOG_N = 20


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    d = tmp_path_factory.mktemp("ref_calls")
    assert os.path.exists(BIN), "tests/cpp/ref_calls not built (make -C tests/cpp)"
    r = subprocess.run([BIN, str(d), str(LOG_N), "100"], capture_output=True, text=True, timeout=600)
    status = (d / "status.txt").read_text() if (d / "status.txt").exists() else ""
    assert r.returncode == 0 and status.startswith("ok"), (r.returncode, status, r.stdout[-2000:], r.stderr[-2000:])
    return d


def _u64(path):
    return np.fromfile(path, dtype=np.uint64)


def _sample(n, count=512):
    return np.unique(np.concatenate([[0, n - 1], np.random.default_rng(7).integers(0, n, count)]))


def test_g1_fixed_base_pre_comp_and_msm(out, orc):
    n = 1 << LOG_N
    k, s = _u64(out / "k_g1.bin"), _u64(out / "s_g1.bin")
    lvl = _u64(out / "lvl0_g1.bin").reshape(n, 12)
    idx = _sample(n)
    want = orc.g1_mul_gen(np.concatenate([k[4 * i: 4 * i + 4] for i in idx])).reshape(-1, 8)
    assert np.array_equal(lvl[idx, :8], want)
    one = orc.g1_generator()[8:12]     # Montgomery one (the generator's Z)
    assert all(np.array_equal(lvl[i, 8:], one) for i in idx)
    assert np.array_equal(_u64(out / "dest_g1.bin"), orc.msm_closed_form(s, k, 1))


def test_g2_fixed_base_pre_comp_and_msm(out, orc):
    n = 1 << LOG_N
    k, s = _u64(out / "k_g2.bin"), _u64(out / "s_g2.bin")
    rec = _u64(out / "lvl0_g2.bin").reshape(n, 2, 12)   # lane records c0 | c1: X.ci | Y.ci | Z.ci
    idx = _sample(n, 128)
    want = orc.g2_mul_gen(np.concatenate([k[4 * i: 4 * i + 4] for i in idx])).reshape(-1, 4, 4)  # x0 x1 y0 y1
    for j, i in enumerate(idx):
        assert np.array_equal(rec[i, 0, 0:4], want[j, 0]) and np.array_equal(rec[i, 1, 0:4], want[j, 1])
        assert np.array_equal(rec[i, 0, 4:8], want[j, 2]) and np.array_equal(rec[i, 1, 4:8], want[j, 3])
    assert np.array_equal(_u64(out / "dest_g2.bin"), orc.msm_closed_form(s, k, 2))


def test_ntt_setup_table(out, orc):
    assert np.array_equal(_u64(out / "ntt_params.bin"), orc.ntt_setup(1 << 10))


def test_matmul_spmat(out, orc):
    for which, name in enumerate("abc"):
        rp, col, _ = orc.matmul_csr(3, which)
        assert np.array_equal(_u64(out / f"spmat_{name}_col.bin"), col), name
        assert np.array_equal(_u64(out / f"spmat_{name}_rp.bin"), rp), name


def test_groth16_bn254_matmul_p100(out, orc):
    p, seed, wseed = map(int, (out / "g16.txt").read_text().split())
    assert p == 100
    o = orc.Groth16(seed, matmul_p=p)
    zz = o.witness(wseed)
    proof = _u64(out / "g16_proof.bin")
    assert np.array_equal(proof, o.prove_trapdoor(zz, o.rs(wseed)))
    assert o.verify(zz[: 4 * (o.num_public + 1)], proof)
