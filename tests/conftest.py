import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cudabifurcation3.0_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    import orclib
    orclib.lib()
    return orclib


@pytest.fixture(scope="session")
def bh():
    """The product library; GPU tests fail (not skip) when it cannot run."""
    import bifhip
    bifhip.load()
    return bifhip


@pytest.fixture(scope="session")
def dev(bh):
    """Device 0 through the library's own HIP runtime.  GPU tests never
    import torch: torch bundles a second HIP runtime, and two runtimes in one
    process do not share the device."""
    assert bh.device_count() > 0, "GPU test collected on a host without a GPU"
    return 0


@pytest.fixture(autouse=True)
def _fresh_context(request):
    """Reset the library's knobs before every GPU test."""
    if request.node.get_closest_marker("gpu") is not None:
        import bifhip
        bifhip.reset_context()
    yield
