"""Test configuration: import paths, the ``gpu`` marker, and built-library fixtures.

Tests marked ``gpu`` run only on a machine with a HIP device (MI355X); everything else runs on
the CPU and never launches a kernel.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "restir-vulkan_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls through the C ABI")


@pytest.fixture(scope="session")
def product_lib():
    from restir_amd import abi
    return abi.load_library()


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.lib()
