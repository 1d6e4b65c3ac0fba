import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The loaded HIP library binding; GPU tests fail loudly without it."""
    if not gpu_available():
        pytest.fail("GPU test collected on a machine without a ROCm GPU")
    from emsim_amd.backend import hip_ext
    return hip_ext


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    if os.environ.get("EMSIM_REQUIRE_GPU"):
        return
    skip = pytest.mark.skip(reason="no ROCm GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
