import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "low-fidelity-dynamic-models_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP solver)")


@pytest.fixture(scope="session")
def golden18():
    with np.load(os.path.join(GOLDEN, "d18_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden11():
    with np.load(os.path.join(GOLDEN, "d11_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden24():
    with np.load(os.path.join(GOLDEN, "d24_golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
