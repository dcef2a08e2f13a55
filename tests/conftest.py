import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs through libtomocam_hip.so)")


@pytest.fixture(scope="session")
def torch_cuda():
    """torch with a visible GPU; the HIP library must be loadable."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from tomocam_amd import _lib
    _lib.lib()  # fails loudly if the extension is missing
    return torch


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    return float(np.linalg.norm((a - b).ravel()) /
                 max(np.linalg.norm(b.ravel()), 1e-300))


def announce(msg: str) -> None:
    """Prints the case about to run on the GPU (flushed before launch)."""
    print(f"[gpu-case] {msg}", flush=True)
