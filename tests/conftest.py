import os
import sys

import numpy as np
import pytest

import itensor_amd  # noqa: F401  (imports torch before the native libraries)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def torch_dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def ctx(torch_dev):
    import torch

    from itensor_amd.itx import Context

    c = Context(0, torch.cuda.current_stream().cuda_stream)
    yield c
    torch.cuda.synchronize()
    c.close()


class Dev:
    """numpy <-> device helpers (column-major tensors as flat float64 buffers)."""

    def __init__(self, torch_dev):
        import torch

        self.torch = torch
        self.dev = torch_dev

    def put(self, a: np.ndarray):
        flat = np.asarray(a, dtype=np.float64).ravel(order="F")
        return self.torch.from_numpy(flat.copy()).to(self.dev)

    def empty(self, n: int, fill=float("nan")):
        return self.torch.full((max(1, int(n)),), fill, dtype=self.torch.float64, device=self.dev)

    def get(self, t, dims):
        self.torch.cuda.synchronize()
        flat = t.detach().cpu().numpy()[: int(np.prod(dims))]
        return flat.reshape(dims, order="F")


@pytest.fixture(scope="session")
def dev(torch_dev):
    return Dev(torch_dev)


def relerr(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = max(np.max(np.abs(ref)) if ref.size else 0.0, 1e-300)
    return float(np.max(np.abs(x - ref)) / den) if ref.size else 0.0
