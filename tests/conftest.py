import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def gpu():
    """Skips on hosts without a GPU; on a GPU host the HIP library must load."""
    import torch  # noqa: F401  (initialises the ROCm runtime like the product does)

    from pyhea_amd import _native

    if _native.device_count() < 1:
        pytest.skip("no GPU visible")
    return _native
