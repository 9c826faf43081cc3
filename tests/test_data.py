"""Synthetic workload generators (dsgdpp_amd/data.py): distinct cells,
disjoint train/test, capped Zipf degrees (water-filling)."""
import numpy as np

from dsgdpp_amd import data


def test_zipf_degrees_capped_and_sum():
    deg = data.zipf_degrees(1000, 50000, 1.0, 120)
    assert abs(deg.sum() - 50000) < 1e-6 * 50000
    assert deg.max() <= 120 + 1e-9
    assert np.all(np.diff(deg) <= 1e-9)                 # non-increasing in rank
    free = deg[deg < 120]
    k = np.arange(1, 1001)[deg < 120]
    assert np.allclose(free * k, free[0] * k[0])        # C / k beyond the cap


def test_uniform_problem_distinct_disjoint():
    p = data.uniform_problem(300, 400, 5000, 500, rank=4, seed=3)
    tr = set(zip(p["train"][0].tolist(), p["train"][1].tolist()))
    te = set(zip(p["test"][0].tolist(), p["test"][1].tolist()))
    assert len(tr) == 5000 and len(te) == 500 and not (tr & te)
    assert p["W0"].shape == (300, 4) and p["H0"].shape == (4, 400)
    assert np.abs(p["W0"]).max() <= 0.5


def test_zipf_problem_values_and_skew():
    p = data.zipf_problem(2000, 300, 40000, 1000, rank=8, truth_rank=4, seed=4)
    x = p["train"][2]
    assert set(np.unique(x)) <= {1.0, 2.0, 3.0, 4.0, 5.0}
    cells = p["train"][0].astype(np.int64) * 300 + p["train"][1]
    assert np.unique(cells).size == cells.size
    cdeg = np.sort(np.bincount(p["train"][1], minlength=300))[::-1]
    assert cdeg[0] > 10 * np.median(cdeg)               # heavy head
