"""pytest configuration: the `gpu` marker and shared fixtures.

CPU tests (-m "not gpu") exercise the oracle against its pins, the host
logic and the library's exported ABI.  GPU tests (-m gpu) call the HIP
library through the C ABI and compare with the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle.loader import OracleHelpers, load_oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _load_oracle():
    return load_oracle()


@pytest.fixture(scope="session")
def oracle():
    return _load_oracle()


@pytest.fixture(scope="session")
def ohelp(oracle):
    return OracleHelpers(oracle)


@pytest.fixture(scope="session")
def dev():
    import heterosync_amd as hs
    d = hs.Device(0)
    yield d
    d.close()
