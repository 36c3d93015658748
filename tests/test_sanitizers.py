"""The CPU oracle and the library's host builders (tree, alias tables, light clusters) under
AddressSanitizer and UndefinedBehaviorSanitizer: oracle/sanitize_check.cpp, built by
`make -C oracle sanitize`, runs them on seeded cases (including the dominant-light and zero-power
alias tables and a frame of every pass) and compares the host builders with the oracle. Any
out-of-bounds access or undefined behaviour aborts it."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parent.parent / "oracle"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with ASan/UBSan")
def test_oracle_and_host_builders_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(ORACLE / "_san" / "sanitize_check")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "clean under ASan/UBSan" in r.stdout
