"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

tests/cpp/host_sanitize.cpp runs the product's host arithmetic
(liblightzk_amd/csrc/host_ec.hpp), its seeded input streams (synth.cpp), the
proving-key file checks (keyfile_format.hpp: header, size, R1CS, trailer)
on a valid file and on each defect the loader rejects, and the oracle's CPU
restatement, built with g++ -fsanitize=address,undefined (no GPU code is
involved). Any sanitizer report or wrong value fails the run."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "sanitize"], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(ROOT, "tests", "cpp", "host_sanitize"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0 and "host_sanitize ok" in r.stdout, (r.returncode, r.stdout[-3000:], r.stderr[-5000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-5000:]
