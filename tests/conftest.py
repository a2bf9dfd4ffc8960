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
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU test collected on a host without a GPU"
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _fresh_context(request):
    """Reset the library's knobs before every GPU test."""
    if request.node.get_closest_marker("gpu") is not None:
        import bifhip
        bifhip.reset_context()
    yield
