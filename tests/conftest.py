import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the gfx950 kernels)")


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def relmse(img, ref, eps=1e-2):
    """mean over pixels and channels of (x - y)^2 / (y^2 + eps)."""
    import numpy as np
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


def film_image(film):
    import numpy as np
    return film[:3] / np.maximum(film[3], 1e-12)
