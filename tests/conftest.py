"""Test configuration: the ``gpu`` marker and import paths.

Tests marked ``gpu`` need an AMD GPU (they call the HIP library through its C
ABI); everything else runs on the CPU (the oracle, host-side logic and the
library's exported interface).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "optical-flow-filter_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ffo():
    import ffo as mod
    return mod
