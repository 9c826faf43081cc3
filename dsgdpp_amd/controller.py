"""Learned step-size controller for DSGD++ epochs.

The reference chooses each epoch's step size with a decay functor
(mf/sgd/decay/decay_bolddriver.h:59-67: BoldDriver grows the step by 1.05
after an improving epoch and halves it otherwise).  This module adds a
learned choice:

* state   -- the last 8 epochs' training-loss ratios and the current step
             size (9 features, see ``features``);
* network -- a 9 -> 64 -> 64 -> 1 GELU perceptron in FP32, evaluated and
             trained on the GPU by the library's own kernels
             (``dsgdpp_mlp_*`` in include/dsgdpp.h, dsgdpp_amd/csrc/mlp.hip);
* action  -- the network's output is the log of the step multiplier for the
             next epoch, clamped to [1/2, 2];
* fallback -- BoldDriver's rejection rule stays: after an epoch that did not
             lower the training loss the step is halved, whatever the network
             says.

Training targets come from DSGD++ itself: along a trajectory, every epoch is
run from the same factors with several step multipliers, the training losses
are fitted by a parabola in log(multiplier), and the minimiser (clamped to
the action range) is the target; the trajectory continues from the best
candidate (or, for a second family of trajectories, with BoldDriver's own
choice, so the states BoldDriver visits are covered too).
"""
import ctypes
import math
import os
import sys

import numpy as np

from . import _lib
from ._lib import TEST, TRAIN, ptr

N_HIST = 8
N_IN = N_HIST + 1
LOG_MIN, LOG_MAX = math.log(0.5), math.log(2.0)
CANDIDATES = (0.5, 0.7071, 1.0, 1.4142, 2.0)
DEFAULT_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "controller_params.npz")


def features(losses, eps):
    """9 features from the training losses so far (losses[0] = before the
    first epoch) and the current step size: the log ratios of the last 8
    consecutive losses (newest last, clipped to [-1, 1], 0 where the history
    is shorter) and log10(eps) / 5."""
    f = np.zeros(N_IN, dtype=np.float32)
    r = []
    for a, b in zip(losses[:-1], losses[1:]):
        if a > 0 and b > 0 and math.isfinite(a) and math.isfinite(b):
            r.append(max(-1.0, min(1.0, math.log(b / a))))
        else:
            r.append(1.0)
    r = r[-N_HIST:]
    f[N_HIST - len(r):N_HIST] = r
    f[N_HIST] = math.log10(max(eps, 1e-30)) / 5.0
    return f


def init_params(seed=0):
    """Scaled uniform initialisation (fan-in), small last layer."""
    rng = np.random.default_rng(seed)
    h = 64

    def u(shape, fan_in, gain=1.0):
        a = gain * math.sqrt(3.0 / fan_in)
        return rng.uniform(-a, a, shape)

    parts = [u((h, N_IN), N_IN), np.zeros(h), u((h, h), h), np.zeros(h), u((h,), h, 0.1), np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


class Mlp:
    """The controller network on one GPU (C ABI ``dsgdpp_mlp_*``)."""

    def __init__(self, device=0, params=None):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.dsgdpp_mlp_create(int(device), ctypes.byref(h))
        if rc != _lib.OK:
            raise RuntimeError(f"dsgdpp_mlp_create failed ({rc}): {self._lib.dsgdpp_mlp_create_error().decode()}")
        self._h = h
        self.n_params = self._lib.dsgdpp_mlp_num_params()
        if params is not None:
            self.set_params(params)

    def _check(self, rc, what):
        if rc != _lib.OK:
            raise RuntimeError(f"{what} failed ({rc}): {self._lib.dsgdpp_mlp_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self._lib.dsgdpp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        # At interpreter exit the HIP runtime and profiler hooks may already be
        # torn down: release nothing then (the process is ending anyway).
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:   # noqa: BLE001
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        self._check(self._lib.dsgdpp_mlp_set_params(self._h, ptr(p, ctypes.c_float)), "set_params")

    def get_params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._lib.dsgdpp_mlp_get_params(self._h, ptr(p, ctypes.c_float)), "get_params")
        return p

    def forward(self, X):
        X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float32)
        out = np.zeros(X.shape[0], dtype=np.float32)
        self._check(self._lib.dsgdpp_mlp_forward(self._h, ptr(X, ctypes.c_float), X.shape[0],
                                                 ptr(out, ctypes.c_float)), "forward")
        return out

    def gradient(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self._lib.dsgdpp_mlp_gradient(self._h, ptr(X, ctypes.c_float), ptr(Y, ctypes.c_float),
                                                  X.shape[0], ptr(g, ctypes.c_float), ctypes.byref(loss)),
                    "gradient")
        return loss.value, g

    def train(self, X, Y, steps, lr=1e-3):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        losses = np.zeros(steps, dtype=np.float64)
        self._check(self._lib.dsgdpp_mlp_train(self._h, ptr(X, ctypes.c_float), ptr(Y, ctypes.c_float),
                                               X.shape[0], steps, ctypes.c_float(lr),
                                               ptr(losses, ctypes.c_double)), "train")
        return losses


class LearnedDecay:
    """Step-size rule with the BoldDriver call signature: ``eps = rule(prev,
    cur)`` before each epoch (prev/cur = training loss before/after the last
    epoch, None before the first two calls as in api.BoldDriver)."""

    def __init__(self, eps0, mlp=None, params=None, dec=0.5):
        self.eps = float(eps0)
        self.dec = float(dec)
        self.losses = []
        self.mlp = mlp if mlp is not None else Mlp(params=load(params) if not isinstance(params, np.ndarray)
                                                    else params)
        self.fallbacks = 0

    def __call__(self, prev, cur):
        if cur is not None and (not self.losses or self.losses[-1] != cur):
            if not self.losses and prev is not None:
                self.losses.append(prev)
            self.losses.append(cur)
        if prev is None:
            return self.eps
        if not (prev > cur) or not math.isfinite(cur):   # BoldDriver's rejection rule
            self.eps *= self.dec
            self.fallbacks += 1
            return self.eps
        y = float(self.mlp.forward(features(self.losses, self.eps))[0])
        self.eps *= math.exp(min(LOG_MAX, max(LOG_MIN, y)))
        return self.eps


def save(path, params, meta=None):
    np.savez(path, params=np.asarray(params, dtype=np.float32), **(meta or {}))


def load(path=None):
    with np.load(path or DEFAULT_PARAMS) as z:
        return z["params"].astype(np.float32)


# ---------------------------------------------------------------------------
# training data from DSGD++ trajectories
# ---------------------------------------------------------------------------
def _target(mults, losses):
    """Minimiser of a parabola through the best candidate and its neighbours
    (in log multiplier), clamped to the action range."""
    lm = np.log(np.asarray(mults))
    ls = np.asarray(losses, dtype=np.float64)
    k = int(np.argmin(ls))
    if not np.all(np.isfinite(ls)):
        return float(lm[k])
    lo, hi = max(0, k - 1), min(len(lm) - 1, k + 1)
    if hi - lo < 2:
        lo, hi = (0, 2) if k == 0 else (len(lm) - 3, len(lm) - 1)
    x, y = lm[lo:hi + 1], ls[lo:hi + 1]
    a = np.polyfit(x, y, 2)
    if a[0] <= 0:
        return float(lm[k])
    return float(min(LOG_MAX, max(LOG_MIN, -a[1] / (2 * a[0]))))


def trajectory(problem, update, eps0, epochs, follow="greedy", tasks=4, seed=1, precision="FP32", kernel=0,
               device=0):
    """(X, Y) samples along one trajectory on the GPU.  follow="greedy"
    continues from the best candidate, "bold" from BoldDriver's choice."""
    from . import api
    g = api.DsgdPp(update=update, loss=update, rank=problem["rank"], tasks_per_rank=tasks, sgd_order="WOR",
                   stratum_order="COWOR", seed=seed, precision=precision, kernel=kernel, device=device)
    g.load_coo(TRAIN, problem["m"], problem["n"], *problem["train"])
    g.set_factors(problem["W0"], problem["H0"])
    losses = [g.loss(TRAIN)]
    eps = float(eps0)
    bold = api.BoldDriver(eps0)
    bold(None, losses[0])
    X, Y = [], []
    for _ in range(epochs):
        W, H = g.get_factors()
        cand = []
        for m in CANDIDATES:
            g.set_factors(W, H)
            g.epoch(eps * m)
            lc = g.loss(TRAIN)
            cand.append((lc, m, g.get_factors() if follow == "greedy" else None))
        cl = [c[0] for c in cand]
        X.append(features(losses, eps))
        Y.append(_target(CANDIDATES, cl))
        if follow == "greedy":
            k = int(np.argmin([c if math.isfinite(c) else math.inf for c in cl]))
            g.set_factors(*cand[k][2])
            losses.append(cl[k])
            eps *= CANDIDATES[k]
        else:
            g.set_factors(W, H)
            g.epoch(eps)
            losses.append(g.loss(TRAIN))
            eps = bold(losses[-2], losses[-1])
    g.close()
    return np.array(X, dtype=np.float32), np.array(Y, dtype=np.float32)


def training_problems(seed=0):
    """Small problems of the three families the controller serves: the
    uniform quick-start shape, the same at rank 128, and the Zipf shape."""
    from . import data
    out = []
    for k, (fn, kw, update, eps0) in enumerate((
            (data.quickstart_1m, dict(shrink=20), "Nzsl_L2(1)", 0.001),
            (data.quickstart_1m, dict(shrink=10), "Nzsl_L2(0.1)", 0.01),
            (data.quickstart_1m_r128, dict(shrink=20), "Nzsl_L2(1)", 0.001),
            (data.zipf_100m, dict(shrink=400), "Nzsl_L2(0.05)", 0.01),
            (data.zipf_100m, dict(shrink=200), "Nzsl_L2(0.05)", 0.001))):
        out.append((fn(seed=seed + 11 * k, **kw), update, eps0))
    return out


def build_dataset(epochs=25, seed=0, log=None):
    X, Y = [], []
    for k, (p, update, eps0) in enumerate(training_problems(seed)):
        for follow in ("greedy", "bold"):
            x, y = trajectory(p, update, eps0, epochs, follow=follow, seed=seed + k + 1)
            X.append(x)
            Y.append(y)
            if log:
                log(f"trajectory {k} ({follow}): {len(y)} samples, mean target {float(np.mean(y)):.3f}")
    return np.concatenate(X), np.concatenate(Y)


def fit(X, Y, steps=3000, lr=1e-3, seed=0, device=0):
    net = Mlp(device=device, params=init_params(seed))
    losses = net.train(X, Y, steps, lr)
    return net.get_params(), losses


def main():
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="train the step-size controller on DSGD++ trajectories")
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    ap.add_argument("--epochs", type=int, default=25)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()

    def log(s):
        print(s, file=sys.stderr, flush=True)

    X, Y = build_dataset(a.epochs, a.seed, log)
    params, losses = fit(X, Y, a.steps, seed=a.seed)
    log(f"{len(Y)} samples; training mse {losses[0]:.4f} -> {losses[-1]:.4f}")
    save(a.out, params, dict(n_samples=len(Y), mse_first=losses[0], mse_last=losses[-1]))
    log(f"wrote {a.out}")


if __name__ == "__main__":
    main()
