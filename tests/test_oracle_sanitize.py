"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
driver oracle/sanitize_main.c runs every oracle entry point over edge-case
clouds (grid-aligned edges, z ties, one- and two-point clouds, all leaf sizes,
one and four threads) with every sanitizer report fatal, and cross-checks the
tree stages against brute force."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "_build", "oracle_sanitize")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle sanitize: clean" in r.stdout
    assert "runtime error" not in r.stderr
