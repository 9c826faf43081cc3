"""Batched epochs (one persistent launch over several independent problems)
and the task-count autotuner.  GPU tests: every problem of a batch ends each
epoch bit-identical to the oracle's emulation of its precision (the grouped
kernel's association on the same schedules and visit orders), with and
without the captured graph; the rank-128 lane layout of the grouped kernel
matches the oracle's float emulation bit for bit."""
import numpy as np
import pytest

from dsgdpp_amd import batch as B


def test_plan_tasks_balances_predicted_time():
    rates = {r: {1: 1e7, 2: 1e7, 4: 9e6, 8: 8e6, 16: 6e6, 32: 4e6, 64: 3e6} for r in B.RANK_CLASSES}
    probs = [dict(m=10_000, n=5_000, nnz=8_000_000, rank=16),
             dict(m=10_000, n=5_000, nnz=1_000_000, rank=64),
             dict(m=40, n=30, nnz=600, rank=128)]
    mins = [B.min_tasks(p) for p in probs]
    # blocks below 8000 points (16-bit prep): 8M points need d = 32, 1M points d = 16
    assert mins == [32, 16, 1]
    assert all(B.block_cap(p["nnz"], m) < B.PREP_NARROW for p, m in zip(probs, mins))
    assert B.block_cap(8_000_000, 16) >= B.PREP_NARROW and B.block_cap(1_000_000, 8) >= B.PREP_NARROW
    # the LDS alone: 8M points need d = 16 for the prep state; the wide kernel
    # streams H, so the tiles do not bind, and without it two rank-64 tiles of
    # 5000 / 2d columns need d = 16
    assert B.prep_lds_estimate(10_000, 5_000, 8_000_000, 8) > B.LDS_BYTES
    assert B.prep_lds_estimate(10_000, 5_000, 8_000_000, 16) <= B.LDS_BYTES
    assert B.lds_need(5_000, 8, 64) > B.LDS_BYTES >= B.lds_need(5_000, 8, 64, wide=True)
    assert B.lds_need(5_000, 16, 64) <= B.LDS_BYTES
    # the narrow-prep rule stops at the shape (d <= m, 2d <= n): 600k points on 40 x 30 stop at d = 8
    assert B.min_tasks(dict(m=40, n=30, nnz=600_000, rank=16)) == 8
    d = B.plan_tasks(probs, budget=120, rates=rates)
    assert sum(d) <= 120 and all(a >= b for a, b in zip(d, mins))
    # the big problem is at the largest measured d; the others grow until the shape stops them
    assert d == [32, 32, 8]
    d = B.plan_tasks(probs, budget=60, rates=rates)
    assert d[0] == 32 and d[1] == 16 and sum(d) <= 60   # no room to grow the 1M problem
    pred = [p["nnz"] / (dk * rates[16][dk]) for p, dk in zip(probs, d)]
    assert max(pred) == pred[0]
    assert probs[2]["m"] >= d[2] and probs[2]["n"] >= 2 * d[2]
    assert B.plan_tasks(probs, budget=3, rates=rates) == mins   # never below the LDS minimum
    parts = B.partition([8, 8, 1, 20, 5], budget=21)
    assert sorted(sum(parts, [])) == [0, 1, 2, 3, 4]
    assert all(sum([8, 8, 1, 20, 5][k] for k in q) <= 21 for q in parts)
    # by time: the two slow problems share a launch, the fast ones the other
    tasks, times = [8, 8, 8, 8], [1.0, 9.0, 1.5, 8.0]
    parts = B.partition(tasks, budget=16, times=times)
    assert sorted(parts) == [[0, 2], [1, 3]]
    assert sum(max(times[k] for k in q) for q in parts) == 10.5
    assert sum(max(times[k] for k in q) for q in B.partition(tasks, budget=16)) > 10.5


def test_rank_class():
    assert [B.rank_class(r) for r in (1, 16, 17, 64, 65, 128)] == [16, 16, 32, 64, 128, 128]
    with pytest.raises(ValueError):
        B.rank_class(129)


def test_batch_kernel_streams_when_the_tile_leaves_no_room_for_a_prep():
    """The batch's workgroups take the LDS of its widest tile: small tiles
    keep the tile kernel (0); a tile that leaves a CU less LDS than the
    largest prep workgroup needs switches the batch to streaming H (4).  A
    problem whose tile does not fit at all streams anyway and counts with
    the streaming kernel's LDS."""
    small = [dict(m=2000, n=1000, nnz=20000, rank=16)] * 4
    assert B.batch_kernel(small, [4] * 4) == 0
    # 2 x 300 columns x 128 floats: about 340 KB of tile, never fits -> counts as streaming
    huge = dict(m=4000, n=4 * 300, nnz=20000, rank=128)
    assert B.lds_need(huge["n"], 2, 128) > B.LDS_BYTES
    assert B.batch_kernel(small + [huge], [4] * 4 + [2]) == 0
    # a tile of about 150 KB fits on its own, but leaves no room for the prep beside it
    wide = dict(m=8000, n=2 * 4 * 116, nnz=64000, rank=128)
    t = B.lds_need(wide["n"], 4, 128)
    assert t <= B.LDS_BYTES and t + B.prep_lds_estimate(wide["m"], wide["n"], wide["nnz"], 4) > B.LDS_BYTES
    assert B.batch_kernel(small + [wide], [4] * 4 + [4]) == 4
    # a prep too large for LDS keeps its state in global memory and takes no LDS
    assert B.prep_lds_estimate(400000, 1000, 40000, 4) > B.LDS_BYTES
    assert B.batch_kernel(small + [dict(m=400000, n=1000, nnz=40000, rank=16)], [4] * 5) == 0


def _ctx(p, tasks, seed, lam, precision="FP32", kernel=0):
    from dsgdpp_amd import TRAIN, api
    upd = f"Nzsl_L2({lam})"
    g = api.DsgdPp(update=upd, loss=upd, rank=p["rank"], tasks_per_rank=tasks, seed=seed, precision=precision,
                   kernel=kernel)
    g.load_coo(TRAIN, p["m"], p["n"], *p["train"])
    g.set_factors(p["W0"], p["H0"])
    return g


def _problems():
    from test_gpu_parity import skewed_problem
    from dsgdpp_amd import data
    out = [(data.uniform_problem(300, 200, 6000, 0, rank=16, seed=3), 2, 0.05),
           (skewed_problem(400, 300, 9000, 64, seed=4), 3, 0.1),
           (data.uniform_problem(250, 250, 5000, 0, rank=96, seed=5), 4, 0.02),
           (skewed_problem(300, 200, 7000, 128, seed=6), 2, 0.07)]
    return out


def _oracle(p, tasks, seed, lam, precision="FP32"):
    from oracle import pyoracle as po
    upd = f"Nzsl_L2({lam})"
    o = po.Oracle(update=upd, loss=upd, rank=p["rank"], tasks_per_rank=tasks, seed=seed, threads=4)
    o.load_coo(0, p["m"], p["n"], *p["train"])
    o.set_factors(p["W0"], p["H0"])
    o.set_arith(precision)
    return o


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph,pairs,precision,kernel", [(True, None, "FP32", 0), (False, None, "FP32", 0),
                                                              (True, "1", "FP32", 0), (True, None, "BF16", 0),
                                                              (True, None, "FP32", 4), (True, None, "BF16", 4)])
def test_batch_bitexact_vs_oracle(monkeypatch, use_graph, pairs, precision, kernel):
    """Every problem of a batch ends each epoch bit-identical to the oracle's
    emulation of the same precision (same schedules and visit orders, the
    grouped kernel's association), with and without the captured graph.
    pairs "1": every eligible block (FP32, rank <= 64) in two-slot groups.
    kernel 4: every context streams H (the batch's choice when its widest
    tile would leave no LDS for the look-ahead preps beside the launch)."""
    print(f"case: batch of 4 (ranks 16/64/96/128) {precision}, graph={use_graph}, pairs={pairs}, kernel={kernel}",
          flush=True)
    if pairs is not None:
        monkeypatch.setenv("DSGDPP_PAIRS", pairs)
    probs = _problems()
    batch_ctx = [_ctx(p, d, seed=10 + k, lam=lam, precision=precision, kernel=kernel)
                 for k, (p, d, lam) in enumerate(probs)]
    orcs = [_oracle(p, d, seed=10 + k, lam=lam, precision=precision) for k, (p, d, lam) in enumerate(probs)]
    bt = B.DsgdPpBatch(batch_ctx, use_graph=use_graph)
    eps = np.array([0.01, 0.02, 0.005, 0.015])
    for e in range(3):
        st = bt.epoch(eps * (1 + 0.1 * e))
        assert st["launches"] == 1 and st["updates"] == sum(len(p[0]["train"][2]) for p in probs)
        for k, o in enumerate(orcs):
            o.epoch(eps[k] * (1 + 0.1 * e))
        for k in range(len(probs)):
            Wb, Hb = batch_ctx[k].get_factors()
            Wo, Ho = orcs[k].get_factors()
            np.testing.assert_array_equal(Wb, Wo, err_msg=f"problem {k} epoch {e}")
            np.testing.assert_array_equal(Hb, Ho, err_msg=f"problem {k} epoch {e}")
            lb, lo = batch_ctx[k].loss(0), orcs[k].loss(0)
            assert abs(lb - lo) <= 1e-10 * abs(lo), (k, e, lb, lo)
    # all contexts' losses in one call: each context's own loss, bit for bit
    lb = bt.loss(0)
    assert lb.tolist() == [g.loss(0) for g in batch_ctx]
    with pytest.raises(RuntimeError, match="no test matrix"):
        bt.loss(1)
    # the contexts stay usable on their own after batched epochs
    batch_ctx[0].epoch(0.01)
    orcs[0].epoch(0.01)
    np.testing.assert_array_equal(batch_ctx[0].get_factors()[0], orcs[0].get_factors()[0])
    bt.close()
    for g in batch_ctx:
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("rank,precision", [(96, "FP32"), (128, "FP32"), (128, "BF16"), (65, "FP32")])
def test_grouped_wide_rank_vs_emulation(rank, precision):
    """Ranks 65..128 use two float4s per lane (8-element lane partials)."""
    from test_gpu_parity import make_pair, skewed_problem
    print(f"case: grouped rank {rank} {precision} vs emulation", flush=True)
    p = skewed_problem(200, 160, 5000, rank, seed=rank)
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=3, seed=rank + 1, precision=precision, kernel=3)
    o.set_arith(precision)
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == 3
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")


@pytest.mark.gpu
def test_batch_rejects_mixed_contexts():
    print("case: batch rejects mixed precision / FP64", flush=True)
    from dsgdpp_amd import data
    p = data.uniform_problem(100, 80, 1000, 0, rank=8, seed=1)
    a = _ctx(p, 2, 1, 0.1, "FP32")
    b = _ctx(p, 2, 1, 0.1, "BF16")
    c = _ctx(p, 2, 1, 0.1, "FP64_STRICT")
    with pytest.raises(RuntimeError, match="same device, precision"):
        B.DsgdPpBatch([a, b])
    with pytest.raises(RuntimeError, match="grouped kernel"):
        B.DsgdPpBatch([c])
    with pytest.raises(RuntimeError, match="twice"):
        B.DsgdPpBatch([a, a])
    for g in (a, b, c):
        g.close()


@pytest.mark.gpu
def test_batch_wait_limit_reports_instead_of_hanging(monkeypatch):
    """A batched launch is an ordinary launch (a replayed graph node), so its
    dataflow waits carry a time limit.  Forced from the host to one tick, a
    wait gives up on its first unmet poll: every wave leaves its task (the
    tile kernel's mover and compute wave, the H-streaming task), the epoch
    reports the error, and the next epoch under the default limit is again
    bit-identical to the oracle from freshly set factors."""
    from test_gpu_parity import imbalanced_problem
    print("case: batch wait limit forced to one tick (tile + streaming problems)", flush=True)
    # 1,000-column blocks at rank 32: two H tiles exceed the LDS, so H is streamed
    wide = imbalanced_problem(32, d=4, rows=100, cols=1000, heavy=3000, light=12, seed=1)
    probs = [(wide, 4, 0.05), (imbalanced_problem(64, d=4, seed=2), 4, 0.1),
             (imbalanced_problem(16, d=4, seed=3), 4, 0.07)]
    ctxs = [_ctx(p, d, seed=30 + k, lam=lam) for k, (p, d, lam) in enumerate(probs)]
    orcs = [_oracle(p, d, seed=30 + k, lam=lam) for k, (p, d, lam) in enumerate(probs)]
    assert ctxs[0].grouped_ok()[0]
    bt = B.DsgdPpBatch(ctxs, use_graph=True)
    eps = [0.01, 0.02, 0.015]
    monkeypatch.setenv("DSGDPP_WAIT_LIMIT_MS", "0.00001")
    with pytest.raises(RuntimeError, match="dataflow wait gave up"):
        bt.epoch(eps)
    monkeypatch.delenv("DSGDPP_WAIT_LIMIT_MS")
    for k, (p, _, _) in enumerate(probs):
        orcs[k].epoch(eps[k])                    # the oracle's streams pass the failed epoch too
        ctxs[k].set_factors(p["W0"], p["H0"])
        orcs[k].set_factors(p["W0"], p["H0"])
        orcs[k].set_arith("FP32")                # factors as the GPU stores them
    for e in range(2):
        bt.epoch(eps)
        for k, o in enumerate(orcs):
            o.epoch(eps[k])
            Wb, Hb = ctxs[k].get_factors()
            Wo, Ho = o.get_factors()
            np.testing.assert_array_equal(Wb, Wo, err_msg=f"problem {k} epoch {e}")
            np.testing.assert_array_equal(Hb, Ho, err_msg=f"problem {k} epoch {e}")
    bt.close()
    for g in ctxs:
        g.close()


@pytest.mark.gpu
def test_batch_packs_more_tasks_than_cus():
    """40 problems of 8 tasks: more tasks than the device has CUs, so the
    library queues several tasks on each workgroup of the one launch (every
    workgroup's tasks in one order of the problems, at most one task of a
    problem per workgroup).  Every problem ends each epoch bit-identical to the
    oracle's FP32 emulation; a problem of heavy blocks makes the others wait
    on its column blocks across the queues."""
    from test_gpu_parity import imbalanced_problem
    from dsgdpp_amd import data
    print("case: batch of 40 x 8 tasks (packed onto the CUs' workgroups)", flush=True)
    probs = []
    for k in range(40):
        if k == 7:
            probs.append((imbalanced_problem(16, d=8, rows=40, cols=40, heavy=800, light=10, seed=k), 8, 0.05))
        else:
            p = data.uniform_problem(240 + 8 * k, 160 + 4 * k, 3000 + 50 * k, 0, rank=16 if k % 3 else 64,
                                     seed=100 + k)
            probs.append((p, 8, 0.01 + 0.001 * k))
    ctxs = [_ctx(p, d, seed=50 + k, lam=lam) for k, (p, d, lam) in enumerate(probs)]
    orcs = [_oracle(p, d, seed=50 + k, lam=lam) for k, (p, d, lam) in enumerate(probs)]
    bt = B.DsgdPpBatch(ctxs, use_graph=True)
    eps = np.linspace(0.005, 0.02, len(probs))
    for e in range(2):
        st = bt.epoch(eps)
        assert st["launches"] == 1
        for k, o in enumerate(orcs):
            o.epoch(eps[k])
            Wb, Hb = ctxs[k].get_factors()
            Wo, Ho = o.get_factors()
            np.testing.assert_array_equal(Wb, Wo, err_msg=f"problem {k} epoch {e}")
            np.testing.assert_array_equal(Hb, Ho, err_msg=f"problem {k} epoch {e}")
    bt.close()
    for g in ctxs:
        g.close()
