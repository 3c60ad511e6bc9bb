"""pytest configuration: the `gpu` marker and import paths.

Tests marked `gpu` need an MI355X (HIP device) and libdfx.so built in-tree;
everything else runs on the CPU (the oracle, host logic, ABI export checks).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "dft-efe_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (HIP) and libdfx.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    import dfx
    c = dfx.Context(0)
    yield c
    c.close()
