import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "cl-kmeans_amd"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """A clk_context on device 0.  Fails (does not skip) without a GPU."""
    import clkmeans
    from clkmeans import hip
    n = hip.device_count()
    assert n >= 1, "no HIP device visible: the GPU tests need an MI355X"
    ctx = clkmeans.Context(0)
    yield ctx
    ctx.close()
