import os
import sys

import pytest

# Every device buffer the library hands out is filled with 0xFF bytes, so no
# kernel can pass by relying on memory the driver happened to clear.
os.environ.setdefault("RVSP_POOL_POISON", "1")

# One HIP runtime per process: torch brings its own, under the same soname as
# ROCm's. When a GPU is present, torch touches it before any test module
# loads the library, so both bind to the same runtime (the GPU tests that
# build operands with torch need that; the library works on either).
try:
    import torch
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda")
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "rv-sparse_amd"), ROOT, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size configuration, minutes on CPU")


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when a GPU test runs without a usable device: a GPU
    test must exercise the HIP library, never pass silently."""
    import rvsp
    n = rvsp.device_count()
    assert n >= 1, "no usable gfx950 device; the HIP library cannot run"
    return n
