"""Shared pytest setup: import paths and the `gpu` marker.

`-m "not gpu"` runs on any machine (oracle, host logic, ABI exports);
`-m gpu` needs a HIP device and calls libmsj.so through its C ABI.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "self-join-mistic_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libmsj.so")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle  # noqa: WPS433 - test infrastructure
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def msj_mod():
    import msj  # noqa: WPS433
    msj.lib()
    return msj
