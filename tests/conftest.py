import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "mex-cuda-ct_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP operators)")
    config.addinivalue_line("markers", "slow: full-size workload checks")


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")


def rel_errors(a, b):
    """(max|a-b|/max|b|, ||a-b||/||b||) in float64."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = a - b
    mb = max(np.abs(b).max(initial=0.0), 1e-30)
    nb = max(np.linalg.norm(b.ravel()), 1e-30)
    return float(np.abs(d).max(initial=0.0) / mb), float(np.linalg.norm(d.ravel()) / nb)


def rel_rmse(a, ref):
    a = np.asarray(a, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))


# Parity bar for plain FP32 against the oracle (FP and BP).
REL_MAX = 1e-4
REL_L2 = 1e-6
