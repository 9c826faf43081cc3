"""Step-size controller: the numpy restatement of the network pinned by finite
differences, the feature/decay logic on the host, and (GPU) the HIP forward
pass, gradient and training against that restatement, plus the gain of the
trained controller over plain BoldDriver on a problem it was not trained on."""
import math
import os

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from dsgdpp_amd import controller as C
from dsgdpp_amd import TEST, TRAIN


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, mo.N_IN)).astype(np.float32)
    Y = np.tanh(X[:, 0] - 0.5 * X[:, 3] + 0.25 * X[:, 8]).astype(np.float32)
    return X, Y


def test_oracle_gradient_matches_finite_differences():
    X, Y = _batch(7, 1)
    p = C.init_params(3).astype(np.float64)
    p[mo.N_PARAMS - 65:] += np.random.default_rng(4).normal(scale=0.3, size=65)   # non-trivial last layer
    loss, g = mo.loss_and_grad(p, X, Y)
    idx = list(range(0, mo.N_PARAMS, 97)) + [mo.N_PARAMS - 1, mo.N_PARAMS - 2]
    fd = mo.finite_difference(p, X, Y, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-9)
    assert loss == pytest.approx(float(np.mean((mo.forward(p, X) - Y) ** 2)))


def test_oracle_gelu_special_values():
    z = np.array([0.0, 1.0, -1.0, 3.0])
    # GELU(z) = z Phi(z); Phi(1) = 0.8413447460685429
    np.testing.assert_allclose(mo.gelu(z), [0.0, 0.8413447460685429, -0.15865525393145707, 3 * 0.9986501019683699],
                               rtol=1e-12)
    h = 1e-6
    np.testing.assert_allclose(mo.gelu_grad(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h), rtol=1e-7)


def test_features():
    f = C.features([100.0], 1e-3)
    assert f.shape == (9,) and np.all(f[:8] == 0) and f[8] == pytest.approx(-0.6)
    losses = [100.0, 50.0, 40.0, 41.0]
    f = C.features(losses, 0.01)
    np.testing.assert_allclose(f[5:8], [math.log(0.5), math.log(0.8), math.log(41 / 40)], rtol=1e-6)
    assert np.all(f[:5] == 0) and f[8] == pytest.approx(-0.4)
    f = C.features([1.0] + [math.e ** (k + 2) for k in range(12)], 1.0)
    assert np.all(f[:8] == 1.0) and f[8] == 0.0   # clipped ratios, last 8 only


class _StubNet:
    def __init__(self, y):
        self.y = y
        self.calls = 0

    def forward(self, x):
        self.calls += 1
        return np.array([self.y], dtype=np.float32)


def test_learned_decay_fallback_and_clamp():
    d = C.LearnedDecay(0.01, mlp=_StubNet(5.0))
    assert d(None, 10.0) == 0.01
    assert d(10.0, 8.0) == pytest.approx(0.02)          # improvement: clamped to x2
    assert d(8.0, 9.0) == pytest.approx(0.01)           # no improvement: BoldDriver halves
    assert d.fallbacks == 1 and d.mlp.calls == 1
    d = C.LearnedDecay(0.01, mlp=_StubNet(math.log(1.25)))
    d(None, 10.0)
    assert d(10.0, 9.0) == pytest.approx(0.0125)
    assert d(9.0, 9.0) == pytest.approx(0.00625)        # ties count as no improvement
    assert d.losses == [10.0, 9.0]


def test_parabola_target():
    m = np.array(C.CANDIDATES)
    best = 1.3
    losses = (np.log(m) - math.log(best)) ** 2 + 2.0
    assert C._target(C.CANDIDATES, losses) == pytest.approx(math.log(best), abs=1e-9)
    assert C._target(C.CANDIDATES, -np.log(m)) == pytest.approx(math.log(2.0))   # monotone: edge


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    print("case: controller forward/gradient vs numpy", flush=True)
    X, Y = _batch(3000, 7)   # two gradient chunks
    p = C.init_params(5)
    p[-65:] += np.random.default_rng(6).normal(scale=0.3, size=65).astype(np.float32)
    net = C.Mlp(params=p)
    out = net.forward(X)
    np.testing.assert_allclose(out, mo.forward(p, X), rtol=1e-4, atol=1e-5)
    loss, g = net.gradient(X, Y)
    lo, go = mo.loss_and_grad(p, X, Y)
    assert loss == pytest.approx(lo, rel=1e-4)
    np.testing.assert_allclose(g, go, rtol=1e-3, atol=1e-5 * np.abs(go).max())
    _, g2 = net.gradient(X, Y)
    np.testing.assert_array_equal(g, g2)   # fixed summation order


@pytest.mark.gpu
def test_gpu_adam_training_matches_numpy_and_fits():
    print("case: controller Adam steps vs numpy, then fit", flush=True)
    X, Y = _batch(512, 9)
    p0 = C.init_params(1)
    net = C.Mlp(params=p0)
    losses = net.train(X, Y, 3, lr=1e-2)
    # numpy Adam on the oracle gradient
    p = p0.astype(np.float64)
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    ref = []
    for t in range(1, 4):
        l, g = mo.loss_and_grad(p, X, Y)
        ref.append(l)
        m1 = 0.9 * m1 + 0.1 * g
        m2 = 0.999 * m2 + 0.001 * g * g
        p = p - 1e-2 * (m1 / (1 - 0.9 ** t)) / (np.sqrt(m2 / (1 - 0.999 ** t)) + 1e-8)
    np.testing.assert_allclose(losses, ref, rtol=1e-3)
    np.testing.assert_allclose(net.get_params(), p, rtol=1e-3, atol=1e-5)
    more = net.train(X, Y, 400, lr=3e-3)
    assert more[-1] < 0.1 * losses[0]


@pytest.mark.gpu
def test_learned_controller_beats_bolddriver():
    """The committed controller on a Zipf-shaped problem with a seed and size
    it was not trained on (4,801 x 1,777, 1M nonzeros, rank 50,
    Nzsl_L2(0.05)): after the same number of epochs from the same start
    (eps0 = 0.01, the BoldDriver start of the 100M-point workload) its test
    RMSE, the criterion the controller is judged by, is lower than
    BoldDriver's (both training losses are printed)."""
    print("case: learned controller vs BoldDriver", flush=True)
    from dsgdpp_amd import api, data
    assert os.path.exists(C.DEFAULT_PARAMS), "controller_params.npz is missing"
    p = data.zipf_100m(seed=97, shrink=100)
    res = {}
    for name in ("bold", "learned"):
        g = api.DsgdPp(update="Nzsl_L2(0.05)", loss="Nzsl_L2(0.05)", rank=p["rank"], tasks_per_rank=8,
                       seed=5, precision="FP32")
        g.load_coo(TRAIN, p["m"], p["n"], *p["train"])
        g.load_coo(TEST, p["m"], p["n"], *p["test"])
        g.set_factors(p["W0"], p["H0"])
        rule = api.BoldDriver(0.01) if name == "bold" else C.LearnedDecay(0.01)
        prev, cur = None, g.loss(TRAIN)
        for _ in range(15):
            g.epoch(rule(prev, cur))
            prev, cur = cur, g.loss(TRAIN)
        res[name] = (cur, math.sqrt(g.loss(TEST) / len(p["test"][2])))
        g.close()
    print(f"15 epochs: BoldDriver loss {res['bold'][0]:.6g} test rmse {res['bold'][1]:.4f}; "
          f"learned loss {res['learned'][0]:.6g} test rmse {res['learned'][1]:.4f}", flush=True)
    assert res["learned"][1] < res["bold"][1]


class _FakeCtx:
    """Stand-in for api.DsgdPp: the 'factors' are a scalar distance d to the
    optimum; an epoch with step eps multiplies d by |1 - eps/0.01| (best step
    0.01), and the training loss is 1 + d^2."""

    def __init__(self, **kw):
        self.d = None

    def load_coo(self, *a):
        pass

    def set_factors(self, W, H):
        self.d = float(W[0, 0])

    def get_factors(self):
        return np.array([[self.d]]), np.array([[0.0]])

    def epoch(self, eps):
        self.d *= abs(1.0 - eps / 0.01)

    def loss(self, which):
        return 1.0 + self.d * self.d

    def close(self):
        pass


def test_trajectory_targets_on_a_known_landscape(monkeypatch):
    from dsgdpp_amd import api
    monkeypatch.setattr(api, "DsgdPp", _FakeCtx)
    p = dict(m=1, n=1, rank=1, train=(np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1)),
             W0=np.array([[1.0]]), H0=np.array([[0.0]]))
    X, Y = C.trajectory(p, "Nzsl", 0.001, 6, follow="greedy")
    assert X.shape == (6, 9) and Y.shape == (6,)
    assert Y[0] == pytest.approx(math.log(2.0))      # far below the best step: grow as fast as allowed
    assert np.all(Y <= math.log(2.0) + 1e-12) and np.all(Y >= math.log(0.5) - 1e-12)
    Xb, Yb = C.trajectory(p, "Nzsl", 0.001, 6, follow="bold")
    np.testing.assert_allclose(Xb[:, 8], np.log10(0.001 * 1.05 ** np.arange(6)) / 5, rtol=1e-5)
