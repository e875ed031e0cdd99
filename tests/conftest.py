import os
import sys

import numpy as np
import pytest

# One HIP runtime per test process: PyTorch bundles its own libamdhip64, and the
# library binds to whichever copy is loaded first.  Importing torch before any
# test loads the library makes both use torch's, so the device-resident autograd
# tests can hand torch GPU tensors to the library.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libfans_gpu.so")


def sphere32():
    """test/microstructures/sphere32.h5 of the reference (zyx on disk) as [x][y][z]."""
    a = np.fromfile(os.path.join(GOLDEN, "sphere32_ms_zyx_u8.bin"), dtype=np.uint8).reshape(32, 32, 32)
    return np.ascontiguousarray(a.transpose(2, 1, 0)).astype(np.uint16)


@pytest.fixture(scope="session")
def sphere32_ms():
    return sphere32()


def gpu_available():
    try:
        import torch  # noqa: F401  (device count without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
