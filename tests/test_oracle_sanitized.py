"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C
oracle sanitize` builds oracle/_san/oracle_sanitize (oracle/sanitize_main.c),
which runs every oracle function on small seeded cases; any out-of-bounds
access, leak or undefined operation ends it with a report and a non-zero
status."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_runs_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-C", ORACLE, "-s", "sanitize"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    proc = subprocess.run([os.path.join(ORACLE, "_san", "oracle_sanitize")], env=env,
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=300)
    assert proc.returncode == 0, proc.stdout
    assert "clean" in proc.stdout
