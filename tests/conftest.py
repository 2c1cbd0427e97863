"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU.

Both the HIP library and the CPU oracle are loaded from the tree; neither
falls back to the other."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_bmx
    oracle_bmx.lib()
    return oracle_bmx


@pytest.fixture(scope="session")
def bmx():
    import boltzmannmfx_amd
    boltzmannmfx_amd.load_library()
    return boltzmannmfx_amd


@pytest.fixture(scope="session")
def gpu(bmx):
    """A HIP device; a GPU test without one fails loudly rather than skipping."""
    ctx = bmx.Context(0)
    ctx.close()
    return True
