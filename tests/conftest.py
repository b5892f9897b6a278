import os
import subprocess
import sys

import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(TESTS)
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    oracle_so = os.path.join(ROOT, "oracle", "libce_oracle.so")
    if not os.path.exists(oracle_so):
        # the oracle is plain C; build it in-tree when absent (seconds)
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=False)


@pytest.fixture(scope="session")
def gpu():
    """Asserts a GPU and the built library are present for `gpu` tests."""
    from cemc_amd import _capi
    _capi.load()
    n = _capi.device_count()
    assert n >= 1, "no GPU visible"
    return n
