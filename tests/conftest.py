import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def tg_problem(Nx=32, hwr=4, jitter=0.3, seed=0, L=1.0):
    """Taylor-Green strengths (closed-form lattice solve) on a jittered lattice"""
    import oracle as O
    h = L / Nx
    x, y = O.lattice(Nx, L=L)
    rng = np.random.default_rng(seed)
    if jitter:
        x = (x + jitter * h * rng.standard_normal(x.size)) % L
        y = (y + jitter * h * rng.standard_normal(y.size)) % L
    w = (O.tg_vorticity(x, y) / O.tg_lambda11(Nx, hwr, L)).astype(np.float32)
    return x.astype(np.float32), y.astype(np.float32), w, hwr * h


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
