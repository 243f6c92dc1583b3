import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

# FP64 parity: builds without an explicit precision run in FP64 (the reference's
# own FP64 parity runs set the same variable, its default being mixed)
os.environ["ASUKA_ERI_MIXED_PRECISION"] = "0"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        O.build()
    return O


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(HERE, "golden", "rys_golden.npz")
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is present; fail loudly if the HIP library is missing."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from asuka_amd import _lib

    _lib.lib()
    return torch
