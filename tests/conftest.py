import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


def random_problem(m, n, nnz, rank, seed, dup=0, values="normal"):
    """Small seeded COO problem (numpy only): distinct uniform cells (+ `dup`
    duplicated entries), values N(0,1) or small integers, W0/H0 U(-0.5,0.5)."""
    rng = np.random.default_rng(seed)
    cells = rng.choice(m * n, size=nnz, replace=False)
    i = (cells // n).astype(np.int32)
    j = (cells % n).astype(np.int32)
    x = rng.normal(size=nnz) if values == "normal" else rng.integers(1, 6, size=nnz).astype(np.float64)
    if dup:
        k = rng.choice(nnz, size=dup, replace=False)
        i = np.concatenate([i, i[k]])
        j = np.concatenate([j, j[k]])
        x = np.concatenate([x, rng.normal(size=dup)])
        p = rng.permutation(i.size)
        i, j, x = i[p], j[p], x[p]
    W = rng.uniform(-0.5, 0.5, (m, rank))
    H = rng.uniform(-0.5, 0.5, (rank, n))
    return dict(m=m, n=n, train=(i, j, x), W0=W, H0=H, rank=rank)


@pytest.fixture
def problem_factory():
    return random_problem
