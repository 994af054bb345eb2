"""pytest configuration: the `gpu` marker and shared helpers."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")
STATES4 = np.array([2e6, 1.6e6, 1.3e6, 1e6])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def config2_instances(B, n=128, m=None, seed=0):
    """Config-2 distribution: 4 cell states, r ~ U[1,5] ohm, vin in {0, 0.2} V."""
    m = n if m is None else m
    rng = np.random.default_rng(seed)
    st = rng.integers(0, 4, (B, n, m))
    g = (1.0 / STATES4[st]).astype(np.float32)
    r = rng.uniform(1.0, 5.0, B)
    vin = np.where(rng.random((B, n)) < 0.5, 0.2, 0.0)
    return g, r, vin


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fadesim_amd import _native
    _native.lib()  # fail loudly when the HIP library is missing
    return torch.device("cuda:0")
