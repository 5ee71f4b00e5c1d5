"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds oracle/ads_oracle.c with
oracle/sanitize_driver.c, which runs every oracle function (sweeps, step,
source with and without map factors, both CFL criteria, mass sums,
ADS_solve_1D with every boundary-code pair) on small random problems with
exactly sized arrays; any out-of-bounds access, use after free, leak or
undefined operation aborts the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_asan_ubsan():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc/make not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True,
                   capture_output=True)
    exe = os.path.join(ROOT, "oracle", "lib", "ads_oracle_sanitized")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "sanitize_driver: ok" in r.stdout
    assert "runtime error" not in r.stderr
