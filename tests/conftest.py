import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One libcpv context for the GPU tests of a session."""
    import cpvulkan_amd
    ctx = cpvulkan_amd.Context(0)
    yield ctx
    ctx.close()
