"""CPU-side checks of the C ABI boundary: the library loads without a GPU and
exports every function include/hzk.h declares (no compute calls here)."""
import ctypes
import os

import pytest

import liblightzk_amd
from liblightzk_amd import _lib


def test_library_present_and_exports_header():
    assert os.path.exists(_lib.LIB_PATH), "libhzk.so not built"
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_loader_binds_every_header_function():
    L = _lib.load()
    for n in _lib.header_functions():
        assert getattr(L, n).argtypes is not None, f"{n} has no ctypes signature"


def test_version_string_without_gpu():
    assert b"gfx950" in _lib.load().hzk_version()


def test_context_creation_fails_cleanly_without_gpu():
    import subprocess, sys
    # probe GPU presence in a child so this process never initialises HIP
    code = "import ctypes,sys; sys.path.insert(0, %r); import liblightzk_amd as z\n" \
           "try:\n  z.Context(0)\n  print('gpu')\nexcept z.HzkError as e:\n  print('nogpu')\n" % os.path.dirname(
               os.path.dirname(os.path.abspath(liblightzk_amd.__file__)))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() in ("gpu", "nogpu"), out.stderr


def test_product_generator_matches_oracle(orc):
    """The product's seeded generator and the oracle's produce identical bytes."""
    import numpy as np
    for kind, okind in (("fr", False), ("fr_sparse", True)):
        a = liblightzk_amd.gen_field(5, 3, 5000, kind)
        b = orc.gen_fr(5, 3, 5000, sparse=okind)
        assert np.array_equal(a, b), kind
    assert np.array_equal(liblightzk_amd.gen_field(6, 1, 999, "fq"), orc.gen_fq(6, 1, 999))


def test_builtin_plan_table_matches_json(tmp_path):
    """The compiled-in MSM plan table (tune_table.inc, both regimes) is the
    one gen_table.py produces from the shipped mi355x.json."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    here = os.path.join(root, "liblightzk_amd", "tune")
    out = tmp_path / "t.inc"
    subprocess.run([sys.executable, os.path.join(here, "gen_table.py"), os.path.join(here, "mi355x.json"), str(out)],
                   check=True)
    shipped = open(os.path.join(root, "liblightzk_amd", "csrc", "tune_table.inc")).read()
    assert out.read_text() == shipped
    # two regimes x two groups x 32 size classes
    rows = [ln for ln in shipped.splitlines() if ln.startswith("{{")]
    assert len(rows) == 4 and all(r.count("u}") == 32 for r in rows)


def test_ntt_setup_table_matches_oracle(orc):
    """hzk_ntt_setup (host-side, no GPU) writes the reference's ntt_setup table
    (include/protocol/groth16-operator.cuh:6-32) entry for entry, including
    (4M)^-1 at 6M + 1; the oracle's table is pinned to the reference's root
    tables (tests/test_oracle_ntt_spmv.py)."""
    import numpy as np
    for M in (1, 2, 1 << 5, 1 << 12):
        assert np.array_equal(liblightzk_amd.ntt_setup(M), orc.ntt_setup(M)), M
    for bad in (0, 3, 1 << 27):
        with pytest.raises(liblightzk_amd.HzkError):
            liblightzk_amd.ntt_setup(bad)
