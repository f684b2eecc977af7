"""The oracle's C restatement under AddressSanitizer and
UndefinedBehaviorSanitizer (host code, gcc): oracle/san_check.c calls every
entry point of cpu_ref.c in both precisions on 2D, 3D and degenerate shapes
(extent-1 and extent-2 axes), fixed counts and tolerance stops, every NDF
penalty and the forced-stop hook; the run must finish without a report."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_clean_under_asan_ubsan():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "sanitize"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(ORACLE, "_san", "san_check")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "oracle calls clean" in r.stdout
