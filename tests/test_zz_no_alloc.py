"""dsgdpp_loss_noalloc (the loss without allocating or waiting) against
dsgdpp_loss: the same kernels and grids, so the values must be bitwise
equal; two calls in a row, the second on a smaller input (fewer partial sums
in the shared scratch array, the rest of it holding the first call's); every
precision and loss kind; and the call captured into a HIP graph, replayed
with epochs run eagerly between the replays."""
import numpy as np
import pytest

from conftest import random_problem

pytestmark = pytest.mark.gpu


def _ctx(p, update, precision, tasks=4, seed=3, test=None):
    import dsgdpp_amd as D
    g = D.DsgdPp(update=update, loss=update, rank=p["rank"], tasks_per_rank=tasks, seed=seed, precision=precision)
    g.load_coo(D.TRAIN, p["m"], p["n"], *p["train"])
    if test is not None:
        g.load_coo(D.TEST, p["m"], p["n"], *test)
    g.set_factors(p["W0"], p["H0"])
    return g


def _test_matrix(m, n, k, seed):
    rng = np.random.default_rng(seed)
    cells = rng.choice(m * n, size=k, replace=False)
    return (cells // n).astype(np.int32), (cells % n).astype(np.int32), rng.normal(size=k)


CASES = [  # large first, then smaller inputs
    ("FP32", "Nzsl_L2(0.05)", (3000, 2000, 600_000, 24), 40_000),
    ("FP64_STRICT", "Nzsl_Nzl2(0.1)", (400, 300, 20_000, 8), 2_000),
    ("BF16", "Nzsl", (200, 150, 5_000, 16), 500),
    ("FP64", "Nzsl_L2(1)", (60, 50, 700, 5), 100),
]


def test_loss_noalloc_equals_loss():
    import torch
    print("case: dsgdpp_loss_noalloc vs dsgdpp_loss, 4 contexts large to small", flush=True)
    out = torch.full((2 * len(CASES) * 2,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    k = 0
    ctxs = []
    for precision, update, (m, n, nnz, r), ntest in CASES:
        p = random_problem(m, n, nnz, r, seed=m)
        g = _ctx(p, update, precision, test=_test_matrix(m, n, ntest, seed=n))
        ctxs.append(g)
        for epoch in range(2):
            for which in (0, 1):   # training loss, then the smaller test matrix
                want = g.loss(which)
                g.loss_noalloc(which, out[k].data_ptr(), stream.cuda_stream)
                stream.synchronize()
                got = out[k].item()
                assert got == want, (precision, update, epoch, which, got, want)
                k += 1
            g.epoch(0.01)
    # the scratch keeps the last call's partial sums: a large call after the small ones
    want = ctxs[0].loss(0)
    ctxs[0].loss_noalloc(0, out[0].data_ptr(), stream.cuda_stream)
    stream.synchronize()
    assert out[0].item() == want
    with pytest.raises(RuntimeError, match="which must be TRAIN or TEST"):
        ctxs[0].loss_noalloc(2, out[0].data_ptr(), stream.cuda_stream)
    for g in ctxs:
        g.close()


def test_loss_noalloc_in_captured_graph():
    import torch
    print("case: dsgdpp_loss_noalloc captured into a graph, replayed between epochs", flush=True)
    p = random_problem(2000, 1500, 200_000, 32, seed=7)
    g = _ctx(p, "Nzsl_L2(0.05)", "FP32", test=_test_matrix(2000, 1500, 10_000, seed=8))
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        g.loss_noalloc(0, out[0].data_ptr(), s.cuda_stream)
        g.loss_noalloc(1, out[1].data_ptr(), s.cuda_stream)
    for e in range(4):
        g.epoch(0.01 * (1 + e))      # eager work between replays changes the factors
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert out[0].item() == g.loss(0), e
        assert out[1].item() == g.loss(1), e
    g.close()
