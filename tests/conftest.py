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

# PyTorch before the library: torch carries its own copy of the HIP runtime
# under the system runtime's soname, and whichever is loaded first is the
# one the library binds to. Loading torch first gives the process a single
# runtime, so the differentiable form can pass torch tensors to the library.
import torch  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ffo():
    import ffo as mod
    return mod
