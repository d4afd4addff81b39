"""The FP64 CPU oracle and the library's host code (configuration, .ini
parser, initial states, seed-file resampling, diagnostics, energy.csv,
growth-rate fitters, error paths) behind the C ABI, built with
AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer
(oracle/Makefile `sanitize`) and driven by oracle/sanitize_check.cpp.  The
ThreadSanitizer run steps separate contexts from three threads at once (the
C ABI's promise that contexts are independent); it runs each context's
OpenMP loops on one thread, because GCC's OpenMP runtime is not built for
ThreadSanitizer."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEED = os.path.join(REPO, "tests", "golden", "eigenmode_seed_kz1_a1.00_V0.050_sp20.0.bin")
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "ERROR: LeakSanitizer")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "sanitize"], check=True, timeout=600)
    return os.path.join(REPO, "oracle", "_ref_san")


@pytest.mark.parametrize("kind,threads,omp", [("asan", 2, "2"), ("tsan", 3, "1")])
def test_host_code_under_sanitizers(built, tmp_path, kind, threads, omp):
    env = dict(os.environ, OMP_NUM_THREADS=omp, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(built, "check_" + kind), SEED, str(tmp_path), str(threads)],
                       capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "sanitize_check ok" in out
    for rep in REPORTS:
        assert rep not in out, out[-4000:]
