"""The device alias-table build's kernel body (restir-vulkan_amd/csrc/alias_kernel.h, the code
k_alias_build runs) on the CPU with one thread per lane (tests/cpu_kernels/alias_build_cpu.cpp:
1,024 threads, a real workgroup barrier and real 64-lane wave exchanges), under ThreadSanitizer
and under AddressSanitizer + UBSan, on the cases of the GPU tests (one light, equal powers,
uniform, one dominant light paired with every small in turn, a dominant light still large when
the small queue runs dry, zero powers, lognormal, triangle lights); every table must equal the
host build byte for byte.

This build found the cause of the device build's GPU fault: no barrier separated the queue
writes of the normalisation from the first pairing step, so lane 0 could read overA[0] before
the lane that owns the dominant light had stored it (ThreadSanitizer: a data race on the queue;
AddressSanitizer: the stale entry used as a light index faults at scaled[gl])."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent / "cpu_kernels"


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None:
        pytest.skip("needs g++ with ThreadSanitizer and AddressSanitizer")
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return HERE / "_build"


def _run(binary, *args, timeout=600):
    r = subprocess.run([str(binary), *args], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


def test_alias_kernel_body_is_race_free_under_tsan(built):
    rc, out = _run(built / "alias_tsan", "--scale", "50")
    assert "WARNING: ThreadSanitizer" not in out, out[-4000:]
    assert rc == 0 and "all cases equal to the host build" in out, out[-4000:]


def test_alias_kernel_body_under_asan_ubsan(built):
    rc, out = _run(built / "alias_asan", "--scale", "5")
    assert rc == 0 and "all cases equal to the host build" in out, out[-4000:]
    assert "dominant: 1001 lights, equal" in out  # the case whose device run faulted
