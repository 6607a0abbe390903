import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X, gfx950)")
    config.addinivalue_line("markers", "slow: long CPU-side test")


def _gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_available():
        pytest.skip("no GPU visible")
    import torch

    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def ngw_ctx(gpu):
    from nanogw_amd.lanczos import IsdfContext

    ctx = IsdfContext(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def mini():
    """Small synthetic molecule (fast to generate, exact oracle runs in ms)."""
    from nanogw_amd import synth

    return synth.make_molecule(4.5, 0.5, 4, 12, 40, seed=12345)


@pytest.fixture(scope="session")
def mini_tda():
    from nanogw_amd import synth

    return synth.make_molecule(4.0, 0.5, 3, 9, 30, seed=777)


@pytest.fixture(scope="session")
def c1_inputs():
    path = os.path.join(GOLDEN, "c1_inputs.npz")
    if not os.path.exists(path):
        pytest.skip("c1 fixture missing (tests/golden/make_fixtures.py c1)")
    return dict(np.load(path))


@pytest.fixture(scope="session")
def c1_sigma():
    path = os.path.join(GOLDEN, "c1_sigma.npz")
    if not os.path.exists(path):
        pytest.skip("c1 sigma fixture missing")
    return dict(np.load(path))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
