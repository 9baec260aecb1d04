import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True, scope="session")
def _torch_runtime_first():
    """Initialise torch's HIP runtime before the library's: torch ships its own copy of the HIP
    runtime, and once libpqgpu.so's runtime holds the device, torch's later initialisation finds
    no device.  GPU tests pass torch tensors' device pointers through the C-ABI, so torch must
    come first (as bench.py and smoke() do)."""
    try:
        import torch

        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    except Exception:
        pass
    yield
