"""Synthetic workloads (seeded) for the DSGD++ path.

Each generator returns a dict with m, n, train=(i, j, x), test=(i, j, x),
W0 (m x r), H0 (r x n) as numpy arrays (indices int32, values float64).
Generation runs in torch on the given device (a GPU makes the 100M-nonzero
workload take seconds); results are copied to host memory because the C ABI
takes host COO triples.

Rules (recorded with every result):
  uniform   -- cells uniform over the m x n grid without replacement
               (draw, drop duplicates, redraw the shortfall), random
               train/test split of the distinct cells (disjoint);
  zipf(s)   -- expected row/column degrees deg_k = min(cap, C k^-s) over a
               random rank order of the ids (cap = n for rows, m for columns;
               C by water-filling so the degrees sum to the cell count); each
               cell is drawn as (row ~ deg_row, col ~ deg_col) independently
               (Chung-Lu), duplicates dropped and the shortfall redrawn, so
               saturated rows/columns end below their cap.
"""
import math

import numpy as np


def _torch():
    import torch
    return torch


def _gen(device, seed):
    torch = _torch()
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def zipf_degrees(count, total, s, cap):
    """deg_k = min(cap, C k^-s), sum = total (water-filling on C)."""
    k = np.arange(1, count + 1, dtype=np.float64)
    base = k ** (-float(s))
    capped = np.zeros(count, dtype=bool)
    for _ in range(200):
        free = total - cap * capped.sum()
        C = free / base[~capped].sum()
        new = (C * base) > cap
        if (new == capped).all():
            break
        capped = new | capped
    deg = np.where(capped, float(cap), C * base)
    return deg


def _distinct_cells(device, g, m, n, need, row_cdf=None, col_cdf=None):
    """`need` distinct cell ids (row*n + col), in random order."""
    torch = _torch()
    have = torch.empty(0, dtype=torch.int64, device=device)
    for _ in range(64):
        short = need - have.numel()
        if short <= 0:
            break
        draw = int(short * 1.05) + 1024
        if row_cdf is None:
            rows = torch.randint(0, m, (draw,), generator=g, device=device, dtype=torch.int64)
            cols = torch.randint(0, n, (draw,), generator=g, device=device, dtype=torch.int64)
        else:
            u = torch.rand(draw, generator=g, device=device, dtype=torch.float64)
            rows = torch.searchsorted(row_cdf, u).clamp_(max=m - 1)
            u = torch.rand(draw, generator=g, device=device, dtype=torch.float64)
            cols = torch.searchsorted(col_cdf, u).clamp_(max=n - 1)
        cells = rows * n + cols
        have = torch.unique(torch.cat([have, cells]))
    if have.numel() < need:
        raise RuntimeError("could not draw enough distinct cells")
    perm = torch.randperm(have.numel(), generator=g, device=device)
    return have[perm[:need]]


def _split(cells, n, ntrain):
    torch = _torch()
    rows = (cells // n).to(torch.int32)
    cols = (cells % n).to(torch.int32)
    return rows[:ntrain], cols[:ntrain], rows[ntrain:], cols[ntrain:]


def _dots(Wt, Ht, rows, cols, chunk=1 << 23):
    """W*_i . H*_j for every (i, j), chunked."""
    torch = _torch()
    out = torch.empty(rows.numel(), dtype=torch.float64, device=rows.device)
    for a in range(0, rows.numel(), chunk):
        b = min(rows.numel(), a + chunk)
        out[a:b] = (Wt[rows[a:b].long()] * Ht[cols[a:b].long()]).sum(1)
    return out


def _host(*ts):
    return tuple(t.cpu().numpy() for t in ts)


def uniform_problem(m, n, nnz_train, nnz_test, rank, truth_rank=None, noise=0.1, init=0.5,
                    seed=1, device="cpu"):
    """Uniform family (the reference quick-start shape): cells uniform without
    replacement, x = W*_i.H*_j + N(0, noise^2),
    W*, H* iid N(0,1), W0/H0 iid U(-init, init)."""
    torch = _torch()
    g = _gen(device, seed)
    tr = truth_rank or rank
    cells = _distinct_cells(device, g, m, n, nnz_train + nnz_test)
    ri, ci, rt, ct = _split(cells, n, nnz_train)
    Ws = torch.randn(m, tr, generator=g, device=device, dtype=torch.float64)
    Hs = torch.randn(n, tr, generator=g, device=device, dtype=torch.float64)
    x = _dots(Ws, Hs, ri, ci) + noise * torch.randn(ri.numel(), generator=g, device=device, dtype=torch.float64)
    xt = _dots(Ws, Hs, rt, ct) + noise * torch.randn(rt.numel(), generator=g, device=device, dtype=torch.float64)
    W0 = (torch.rand(m, rank, generator=g, device=device, dtype=torch.float64) - 0.5) * (2 * init)
    H0 = (torch.rand(rank, n, generator=g, device=device, dtype=torch.float64) - 0.5) * (2 * init)
    ri, ci, x, rt, ct, xt, W0, H0 = _host(ri, ci, x, rt, ct, xt, W0, H0)
    return dict(m=m, n=n, train=(ri, ci, x), test=(rt, ct, xt), W0=W0, H0=H0, rank=rank,
                rule="uniform", seed=seed)


def zipf_problem(m, n, nnz_train, nnz_test, rank, truth_rank, s=1.0, truth_std=0.3, noise=0.8,
                 offset=3.6, scale=1.0, lo=1.0, hi=5.0, round_values=True, init=0.5, seed=2,
                 device="cpu"):
    """Zipf family: Zipf(s) degrees capped at n (rows) / m (columns);
    x = clip(round(offset + scale * W*_i.H*_j + N(0, noise^2)), lo, hi) with
    W*, H* entries N(0, truth_std^2)."""
    torch = _torch()
    g = _gen(device, seed)
    total = nnz_train + nnz_test
    rdeg = zipf_degrees(m, total, s, n)
    cdeg = zipf_degrees(n, total, s, m)
    rng = np.random.default_rng(seed)
    rdeg = rdeg[rng.permutation(m)]
    cdeg = cdeg[rng.permutation(n)]
    rcdf = torch.tensor(np.cumsum(rdeg) / rdeg.sum(), device=device, dtype=torch.float64)
    ccdf = torch.tensor(np.cumsum(cdeg) / cdeg.sum(), device=device, dtype=torch.float64)
    cells = _distinct_cells(device, g, m, n, total, rcdf, ccdf)
    ri, ci, rt, ct = _split(cells, n, nnz_train)
    Ws = truth_std * torch.randn(m, truth_rank, generator=g, device=device, dtype=torch.float64)
    Hs = truth_std * torch.randn(n, truth_rank, generator=g, device=device, dtype=torch.float64)

    def values(r_, c_):
        v = offset + scale * _dots(Ws, Hs, r_, c_) + noise * torch.randn(
            r_.numel(), generator=g, device=device, dtype=torch.float64)
        if round_values:
            v = torch.round(v)
        return v.clamp_(lo, hi)

    x = values(ri, ci)
    xt = values(rt, ct)
    W0 = (torch.rand(m, rank, generator=g, device=device, dtype=torch.float64) - 0.5) * (2 * init)
    H0 = (torch.rand(rank, n, generator=g, device=device, dtype=torch.float64) - 0.5) * (2 * init)
    ri, ci, x, rt, ct, xt, W0, H0 = _host(ri, ci, x, rt, ct, xt, W0, H0)
    return dict(m=m, n=n, train=(ri, ci, x), test=(rt, ct, xt), W0=W0, H0=H0, rank=rank,
                rule=f"zipf({s})", seed=seed)


# Named workloads: the reference quick-start shape (100k x 100k, 1M uniform
# nonzeros, rank 10), the same at rank 128, and the 100M-nonzero Zipf shape
# (480,189 x 17,770, rank 50); all take a `shrink` factor for small test runs.
def quickstart_1m(seed=1, device="cpu", shrink=1):
    m = n = 100_000 // shrink
    return uniform_problem(m, n, 1_000_000 // shrink, 100_000 // shrink, rank=10, noise=0.1,
                           seed=seed, device=device)


def zipf_100m(seed=2, device="cpu", shrink=1, rank=50):
    return zipf_problem(480_189 // shrink, 17_770 // max(1, int(math.sqrt(shrink))),
                        100_000_000 // shrink, 1_400_000 // shrink, rank=rank, truth_rank=20,
                        s=1.0, truth_std=0.3, noise=0.8, seed=seed, device=device)


def controller_250m(seed=5, device="cpu", shrink=1):
    """The large Zipf shape of the step-size controller study: 1,000,000 x
    625,000, 250M training + 2.5M test nonzeros, Zipf(1.1) degrees, values
    clip(50 + 10 W*_i.H*_j + N(0, 5^2), 0, 100) (not rounded) with rank-30
    N(0, 0.2^2) ground truth, rank 100."""
    return zipf_problem(1_000_000 // shrink, 625_000 // shrink, 250_000_000 // shrink, 2_500_000 // shrink,
                        rank=100, truth_rank=30, s=1.1, truth_std=0.2, noise=5.0, offset=50.0, scale=10.0,
                        lo=0.0, hi=100.0, round_values=False, seed=seed, device=device)


def quickstart_1m_r128(seed=1, device="cpu", shrink=1):
    p = quickstart_1m(seed=seed, device=device, shrink=shrink)
    rng = np.random.default_rng(seed + 100)
    p["W0"] = rng.uniform(-0.5, 0.5, (p["m"], 128))
    p["H0"] = rng.uniform(-0.5, 0.5, (128, p["n"]))
    p["rank"] = 128
    return p
