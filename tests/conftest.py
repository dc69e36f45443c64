import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

PKG = __graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def pkg():
    return PKG


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """torch ships its own HIP runtime next to the one liblqmgpu.so links; torch finds the
    device only when its runtime starts first, so GPU sessions start it before any test (the
    tests that hand torch buffers to the library need both)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch

        torch.cuda.is_available()


@pytest.fixture(scope="session")
def gpu(pkg):
    """The GPU executor; fails loudly (no fallback) if the library or device is missing."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return pkg.api.to_executor("gpu")
