"""The committed fixture tests/golden/tiny_dsgdpp.json (written by
tests/golden/make_golden.py from the oracle) pins schedules, task seeds, visit
orders, losses and factors of a 6 x 8, rank-3, 2 x 4-block problem.  The CPU
test re-runs the oracle; the GPU test runs the HIP path (FP64_STRICT) through
the C ABI and must reproduce every value bit for bit."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_dsgdpp.json")


def load():
    with open(FIX) as f:
        g = json.load(f)
    fx = lambda a: np.array([float.fromhex(v) for v in a])   # noqa: E731
    g["x"] = fx(g["x"])
    g["W0"] = fx(g["W0"]).reshape(g["m"], g["rank"])
    g["H0"] = fx(g["H0"]).reshape(g["rank"], g["n"])
    return g, fx


def _check(h, g, fx):
    l0, w0 = h.loss(0), float.fromhex(g["loss0"])
    assert abs(l0 - w0) <= 1e-13 * abs(w0)   # GPU loss: tree reduction order
    for e, (eps, ep) in enumerate(zip(g["eps"], g["epochs"])):
        h.epoch(eps)
        assert h.schedule().tolist() == ep["schedule"]
        assert [str(int(s)) for s in h.task_seeds()] == ep["seeds"]
        for t in range(2):
            for s in range(4):
                assert h.permutation(t, s).tolist() == ep["perms"][t][s]
        lo = h.loss(0)
        want = float.fromhex(ep["loss"])
        assert abs(lo - want) <= 1e-13 * abs(want)
        if "W" in ep:
            W, H = h.get_factors()
            assert np.array_equal(W.ravel(), fx(ep["W"]))
            assert np.array_equal(H.ravel(), fx(ep["H"]))


def test_oracle_matches_fixture():
    g, fx = load()
    o = po.Oracle(update=g["update"], loss=g["update"], rank=g["rank"], tasks_per_rank=g["d"], seed=g["seed"])
    o.load_coo(0, g["m"], g["n"], np.array(g["i"]), np.array(g["j"]), g["x"])
    o.set_factors(g["W0"], g["H0"])
    o.record(True)
    _check(o, g, fx)


@pytest.mark.gpu
def test_gpu_matches_fixture():
    import dsgdpp_amd as D
    g, fx = load()
    h = D.DsgdPp(update=g["update"], loss=g["update"], rank=g["rank"], tasks_per_rank=g["d"], seed=g["seed"],
                 precision="FP64_STRICT")
    h.load_coo(0, g["m"], g["n"], np.array(g["i"]), np.array(g["j"]), g["x"])
    h.set_factors(g["W0"], g["H0"])
    _check(h, g, fx)
