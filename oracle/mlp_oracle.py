"""ORACLE -- TEST INFRASTRUCTURE ONLY.

CPU restatement (numpy, float64) of the step-size controller network that
dsgdpp_amd/csrc/mlp.hip trains and evaluates on the GPU: a 9 -> 64 -> 64 -> 1
perceptron with exact GELU hidden units, z * Phi(z), and the mean squared
error loss.  The reference has no such network (it extends the decay choices
of mf/sgd/decay/); this file exists so the GPU forward pass and gradient can
be checked against an independent computation, which the tests pin in turn
against central finite differences.

Only tests/ may import this module.  Parameter layout (flat vector):
W1[64x9] (row = hidden unit), b1[64], W2[64x64], b2[64], W3[64], b3[1].
"""
import math

import numpy as np

N_IN, N_H = 9, 64
SIZES = (("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H,)),
         ("b3", (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)

_erf = np.vectorize(math.erf, otypes=[np.float64])


def unpack(p):
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, X):
    """Outputs f(X) for X of shape (n, 9)."""
    q = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ q["W1"].T + q["b1"]
    z2 = gelu(z1) @ q["W2"].T + q["b2"]
    return gelu(z2) @ q["W3"] + q["b3"][0]


def loss_and_grad(p, X, Y):
    """mean((f(X) - Y)^2) and its gradient with respect to the flat parameters."""
    q = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n = X.shape[0]
    z1 = X @ q["W1"].T + q["b1"]
    h1 = gelu(z1)
    z2 = h1 @ q["W2"].T + q["b2"]
    h2 = gelu(z2)
    out = h2 @ q["W3"] + q["b3"][0]
    e = out - Y
    loss = float(np.mean(e * e))
    dout = 2.0 * e / n
    g = {"W3": h2.T @ dout, "b3": np.array([dout.sum()])}
    dz2 = np.outer(dout, q["W3"]) * gelu_grad(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ q["W2"]) * gelu_grad(z1)
    g["W1"] = dz1.T @ X
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def finite_difference(p, X, Y, idx, h=1e-6):
    """Central difference of the loss along the parameters in `idx`."""
    out = []
    for k in idx:
        a = np.array(p, dtype=np.float64)
        b = a.copy()
        a[k] += h
        b[k] -= h
        out.append((loss_and_grad(a, X, Y)[0] - loss_and_grad(b, X, Y)[0]) / (2 * h))
    return np.array(out)
