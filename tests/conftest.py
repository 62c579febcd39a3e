"""Shared pytest setup: repo root on sys.path, the `gpu` marker, fixture paths."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
TABLES = os.path.join(ROOT, "bfpn_gpu_amd", "tables")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The product library on a real device; fails loudly when either is missing."""
    import bfpn_gpu_amd as bg

    lib = bg.lib()  # raises ImportError if the library was not built
    n = bg.device_count()
    assert n > 0, "no HIP device visible: " + lib.bfpn_last_error().decode()
    return bg
