"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

FP64_STRICT: schedules, task seeds and every block's visit order must match
exactly, and the factors must be bitwise identical after every epoch; the
FP64 losses (GPU tree reduction vs the oracle's sequential sums) agree to
1e-12 relative.  Other precisions are checked against the FP64 oracle with the
tolerance written in each test.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import random_problem
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

LOSS_RTOL = 1e-12


def _dsgdpp():
    import dsgdpp_amd
    return dsgdpp_amd


def make_pair(p, update="Nzsl_L2(1)", loss=None, truncate=None, tasks=2, world=1, sgd="WOR",
              stratum="COWOR", seed=7, precision="FP64_STRICT", test=None, kernel=0):
    D = _dsgdpp()
    loss = loss or update
    g = D.DsgdPp(update=update, regularize="None", loss=loss, rank=p["rank"], tasks_per_rank=tasks,
                 world_size=world, sgd_order=sgd, stratum_order=stratum, seed=seed,
                 precision=precision, truncate=truncate, kernel=kernel)
    o = po.Oracle(update=update, regularize="None", loss=loss, truncate=truncate, rank=p["rank"],
                  world_size=world, tasks_per_rank=tasks, sgd_order=sgd, stratum_order=stratum, seed=seed,
                  threads=2)
    for h in (g, o):
        h.load_coo(0, p["m"], p["n"], *p["train"])
        if test is not None:
            h.load_coo(1, p["m"], p["n"], *test)
        h.set_factors(p["W0"], p["H0"])
    o.record(True)
    return g, o


def assert_epoch_parity(g, o, eps, check_perms=True):
    g.epoch(eps)
    o.epoch(eps)
    np.testing.assert_array_equal(g.schedule(), o.schedule())
    np.testing.assert_array_equal(g.task_seeds(), o.task_seeds())
    d = g.d
    if check_perms:
        for t in range(d):
            for s in range(2 * d):
                np.testing.assert_array_equal(g.permutation(t, s), o.permutation(t, s),
                                              err_msg=f"task {t} sub-epoch {s}")
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_array_equal(Hg, Ho)
    lg, lo = g.loss(0), o.loss(0)
    assert abs(lg - lo) <= LOSS_RTOL * abs(lo), (lg, lo)


def test_tiny_strict_bitexact():
    print("case: tiny 12x16 r=3 d=2", flush=True)
    p = random_problem(12, 16, 60, 3, seed=1)
    g, o = make_pair(p, tasks=2)
    for e in range(4):
        assert_epoch_parity(g, o, 0.01 * (1 + e))


@pytest.mark.parametrize("sgd,stratum", list(itertools.product(["SEQ", "WR", "WOR"],
                                                               ["SEQ", "RSEQ", "WR", "WOR", "COWOR"])))
def test_orders_strict(sgd, stratum):
    print(f"case: orders sgd={sgd} stratum={stratum}", flush=True)
    p = random_problem(40, 60, 500, 4, seed=2)
    g, o = make_pair(p, tasks=3, sgd=sgd, stratum=stratum, seed=11)
    for e in range(3):
        assert_epoch_parity(g, o, 0.02)


@pytest.mark.parametrize("update,truncate", [("Nzsl", None), ("Nzsl_L2(0.5)", None),
                                             ("Nzsl_Nzl2(0.05)", None), ("Nzsl_L2(0.5)", "(-0.3,0.3)"),
                                             ("Nzsl", "(-0.2,0.25)")])
def test_update_functors_strict(update, truncate):
    print(f"case: functor {update} truncate={truncate}", flush=True)
    p = random_problem(50, 70, 900, 5, seed=3)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=2, seed=5)
    for e in range(3):
        assert_epoch_parity(g, o, 0.05)


def test_cowor_world_groups_strict():
    print("case: COWOR w=2 t=3", flush=True)
    p = random_problem(60, 80, 1200, 4, seed=4)
    g, o = make_pair(p, tasks=3, world=2, seed=9)
    for e in range(2):
        assert_epoch_parity(g, o, 0.03)


@pytest.mark.parametrize("rank", [1, 63, 64, 65, 128, 130])
def test_ranks_strict(rank):
    print(f"case: rank {rank}", flush=True)
    p = random_problem(30, 40, 300, rank, seed=5)
    g, o = make_pair(p, tasks=2, seed=3)
    for e in range(2):
        assert_epoch_parity(g, o, 0.01)


def test_empty_blocks_and_duplicates_strict():
    print("case: empty blocks + duplicate entries", flush=True)
    p = random_problem(64, 64, 40, 3, seed=6, dup=10)   # d=8: 128 blocks, most empty
    g, o = make_pair(p, tasks=8, seed=13)
    np.testing.assert_array_equal(g.block_nnz(), o.block_nnz())
    for e in range(3):
        assert_epoch_parity(g, o, 0.05)


def test_carry_over_equal_block_sizes_strict():
    """Every block has the same nnz, so each task's permutation carries over
    from block to block (sgd_impl.h:257-262)."""
    print("case: equal-size blocks (permutation carry-over)", flush=True)
    m, n, d = 8, 16, 2
    rng = np.random.default_rng(8)
    i, j = [], []
    for b1 in range(d):
        for b2 in range(2 * d):
            cells = rng.choice(16, size=5, replace=False)
            i += list(b1 * 4 + cells // 4)
            j += list(b2 * 4 + cells % 4)
    p = dict(m=m, n=n, train=(np.array(i), np.array(j), rng.normal(size=len(i))), rank=3,
             W0=rng.uniform(-0.5, 0.5, (m, 3)), H0=rng.uniform(-0.5, 0.5, (3, n)))
    g, o = make_pair(p, tasks=d, seed=17)
    assert (g.block_nnz() == 5).all()
    for e in range(3):
        assert_epoch_parity(g, o, 0.05)


def test_medium_many_blocks_strict():
    print("case: 3000x3000 nnz=60000 r=10 d=16", flush=True)
    p = random_problem(3000, 3000, 60000, 10, seed=9)
    g, o = make_pair(p, tasks=16, seed=21)
    for e in range(2):
        assert_epoch_parity(g, o, 0.005, check_perms=(e == 0))


def test_run_loop_trace_and_test_loss_strict(tmp_path):
    print("case: defaultRunner loop with test matrix", flush=True)
    p = random_problem(200, 300, 6000, 6, seed=10)
    rng = np.random.default_rng(99)
    cells = rng.choice(200 * 300, size=600, replace=False)
    test = ((cells // 300).astype(np.int32), (cells % 300).astype(np.int32), rng.normal(size=600))
    g, o = make_pair(p, tasks=4, seed=31, test=test)
    rg = g.run_native(6, "BoldDriver(0.01)", trace_path=str(tmp_path / "g.R"))
    ro = o.run(6, "BoldDriver(0.01)")
    np.testing.assert_allclose(rg["loss"], ro["loss"], rtol=LOSS_RTOL)
    np.testing.assert_allclose(rg["test_loss"], ro["test_loss"], rtol=LOSS_RTOL)
    np.testing.assert_array_equal(rg["eps"][1:], ro["eps"][1:])
    txt = (tmp_path / "g.R").read_text()
    assert txt.startswith("trace <- list(trace=list(\nlist(it=0, epoch=0, loss=")
    assert txt.count("list(it=") == 7


def test_quickstart_shape_strict_per_epoch_loss():
    """The quick-start shape (reduced 4x) for 8 epochs:
    bit-identical factors and per-epoch loss within 1e-9 relative."""
    print("case: quick-start shape / 4 strict, 8 epochs", flush=True)
    from dsgdpp_amd import data
    p = data.quickstart_1m(shrink=4, seed=3)
    g, o = make_pair(p, tasks=4, seed=1234, test=p["test"])
    rg = g.run_native(8, "BoldDriver(0.001)")
    ro = o.run(8, "BoldDriver(0.001)")
    np.testing.assert_allclose(rg["loss"], ro["loss"], rtol=1e-9)
    np.testing.assert_allclose(rg["test_loss"], ro["test_loss"], rtol=1e-9)
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_array_equal(Hg, Ho)


@pytest.mark.parametrize("precision,rtol", [("FP64", 1e-9), ("FP32", 1e-6), ("BF16", 2e-2)])
def test_precisions_vs_fp64_oracle(precision, rtol):
    """Same schedules/visit orders as the FP64 oracle; per-epoch training loss
    relative error bounded by `rtol` (FP64 lane-parallel dot: 1e-9;
    FP32 factors: 1e-6, observed ~2e-8; BF16 factor storage: 2e-2, observed
    ~6e-4).  A dropped regularization term moves the FP32 loss by ~2e-3."""
    print(f"case: precision {precision}", flush=True)
    p = random_problem(400, 500, 20000, 16, seed=12, values="int")
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=4, seed=41, precision=precision)
    for e in range(5):
        g.epoch(0.01)
        o.epoch(0.01)
        np.testing.assert_array_equal(g.schedule(), o.schedule())
        np.testing.assert_array_equal(g.permutation(1, 3), o.permutation(1, 3))
        lg, lo = g.loss(0), o.loss(0)
        assert abs(lg - lo) <= rtol * abs(lo), (e, lg, lo, abs(lg - lo) / abs(lo))


# ---------------------------------------------------------------------------
# kernel coverage: per-stratum launches (kernel 1) and the dataflow epoch
# (kernel 2, the default for FP64) give bit-identical factors
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kernel", [1, 2])
@pytest.mark.parametrize("sgd,stratum", [("WOR", "COWOR"), ("WR", "WR"), ("SEQ", "WOR")])
def test_kernels_strict(kernel, sgd, stratum):
    print(f"case: kernel {kernel} sgd={sgd} stratum={stratum}", flush=True)
    p = random_problem(80, 120, 3000, 6, seed=21)
    g, o = make_pair(p, tasks=5, sgd=sgd, stratum=stratum, seed=23, kernel=kernel)
    for e in range(3):
        g.epoch(0.02)
        o.epoch(0.02)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo)
        np.testing.assert_array_equal(Hg, Ho)


def skewed_problem(m, n, nnz, rank, seed, hot=0.3):
    """Cells where a few hot rows and columns carry a large share, so groups
    break often and W rows recur within a few groups (forwarding)."""
    rng = np.random.default_rng(seed)
    rw = rng.pareto(1.2, m) + 1
    cw = rng.pareto(1.2, n) + 1
    cells = set()
    while len(cells) < nnz:
        k = nnz - len(cells)
        i = rng.choice(m, size=2 * k, p=rw / rw.sum())
        j = rng.choice(n, size=2 * k, p=cw / cw.sum())
        for a, b in zip(i, j):
            cells.add(int(a) * n + int(b))
            if len(cells) >= nnz:
                break
    cells = np.array(sorted(cells))
    rng.shuffle(cells)
    i = (cells // n).astype(np.int32)
    j = (cells % n).astype(np.int32)
    x = rng.integers(1, 6, size=nnz).astype(np.float64)
    return dict(m=m, n=n, train=(i, j, x), rank=rank, W0=rng.uniform(-0.5, 0.5, (m, rank)),
                H0=rng.uniform(-0.5, 0.5, (rank, n)))


@pytest.mark.parametrize("precision", ["FP32", "BF16"])
@pytest.mark.parametrize("update,truncate", [("Nzsl_L2(0.05)", None), ("Nzsl", None),
                                             ("Nzsl_Nzl2(0.02)", None), ("Nzsl_L2(0.05)", "(-0.4,0.4)")])
def test_grouped_bitexact_vs_emulation(precision, update, truncate):
    """Grouped kernel (kernel 3) against the oracle replaying the same float
    association point by point in visit order: any grouping, forwarding or
    dataflow mistake shows up as a bit difference."""
    print(f"case: grouped {precision} {update} {truncate}", flush=True)
    p = skewed_problem(300, 200, 12000, 50, seed=31)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=4, seed=43, precision=precision, kernel=3)
    o.set_arith(precision)
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == 3
        o.epoch(0.01)
        np.testing.assert_array_equal(g.permutation(2, 5), o.permutation(2, 5))
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
        lg, lo = g.loss(0), o.loss(0)
        assert abs(lg - lo) <= 1e-10 * abs(lo)


@pytest.mark.parametrize("rank,update,truncate", [(50, "Nzsl_L2(0.05)", None), (16, "Nzsl", None),
                                                  (64, "Nzsl_Nzl2(0.02)", None), (37, "Nzsl_L2(0.05)", "(-0.4,0.4)")])
def test_grouped_fp64_bitexact_vs_emulation(rank, update, truncate):
    """FP64 grouped kernel against the oracle replaying its double association
    (4-term FMA chains, the DPP row tree, fused updates) point by point in
    visit order; the epoch loss also stays within 1e-9 of the reference
    arithmetic's."""
    print(f"case: grouped FP64 rank {rank} {update} {truncate}", flush=True)
    p = skewed_problem(300, 200, 12000, rank, seed=31 + rank)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=4, seed=43, precision="FP64", kernel=3)
    _, ref = make_pair(p, update=update, truncate=truncate, tasks=4, seed=43)
    o.set_arith("FP64_GROUPED")
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == 3
        o.epoch(0.01)
        ref.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
        lg, lr = g.loss(0), ref.loss(0)
        assert abs(lg - lr) <= 1e-9 * abs(lr), (e, lg, lr)


@pytest.mark.parametrize("rank,update,truncate", [(50, "Nzsl_L2(0.05)", None), (16, "Nzsl", None),
                                                  (64, "Nzsl_Nzl2(0.02)", None), (37, "Nzsl_L2(0.05)", "(-0.4,0.4)"),
                                                  (100, "Nzsl_L2(0.05)", None), (72, "Nzsl_Nzl2(0.02)", "(-0.4,0.4)")])
def test_grouped_fp64_wide_bitexact_vs_emulation(rank, update, truncate):
    """The FP64 streaming body (H rows from HBM, kernel 4) against the same
    double association as the tile kernel; a shape whose column blocks are
    too wide for the FP64 tile takes it by itself."""
    print(f"case: grouped FP64 wide rank {rank} {update} {truncate}", flush=True)
    p = skewed_problem(300, 200, 12000, rank, seed=131 + rank)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=4, seed=47, precision="FP64", kernel=4)
    o.set_arith("FP64_GROUPED")
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == 4
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
    g.close()
    if rank == 16:   # 2 x 4 blocks of 3,000 columns: the FP64 tile does not fit
        # (kernel 3 asked: on this shape, sparse rows and columns, the automatic
        # choice is the level kernel, tested in test_gpu_level.py)
        rng = np.random.default_rng(9)
        m, n, nnz = 6_000, 12_000, 60_000
        cells = np.unique(rng.integers(0, m * n, size=nnz + nnz // 10))[:nnz]
        rng.shuffle(cells)
        q = dict(m=m, n=n, train=(cells // n, cells % n, rng.normal(size=len(cells))), rank=rank,
                 W0=rng.uniform(-0.5, 0.5, (m, rank)), H0=rng.uniform(-0.5, 0.5, (rank, n)))
        g, o = make_pair(q, update=update, tasks=2, seed=21, precision="FP64", kernel=3)
        o.set_arith("FP64_GROUPED")
        for e in range(2):
            st = g.epoch(0.01)
            o.epoch(0.01)
            assert st["kernel"] == 4
            np.testing.assert_array_equal(g.get_factors()[0], o.get_factors()[0], err_msg=f"wide shape, epoch {e}")
            np.testing.assert_array_equal(g.get_factors()[1], o.get_factors()[1], err_msg=f"wide shape, epoch {e}")
        g.close()


@pytest.mark.parametrize("precision", ["FP32", "BF16"])
@pytest.mark.parametrize("rank,update,truncate", [(50, "Nzsl_L2(0.05)", None), (100, "Nzsl_L2(0.05)", None),
                                                  (16, "Nzsl", None), (72, "Nzsl_Nzl2(0.02)", "(-0.4,0.4)")])
def test_grouped_wide_bitexact_vs_emulation(rank, update, truncate, precision):
    """The grouped kernel with H rows streamed from HBM (kernel 4: H
    prefetch, H forwarding ring, no LDS tile) against the oracle's FP32 or
    BF16 emulation, bitwise: a missed H forwarding (or a BF16 value used
    before its rounding) shows up as a bit difference."""
    print(f"case: grouped wide {precision} rank {rank} {update} {truncate}", flush=True)
    p = skewed_problem(300, 200, 12000, rank, seed=77 + rank)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=4, seed=45, precision=precision, kernel=4)
    o.set_arith(precision)
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == 4
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")


def test_wide_column_blocks_choose_the_streaming_kernel():
    """A 600 x 4,000 problem at d = 2: two LDS tiles of 1,000-column blocks
    (545 KB) cannot be staged, so the automatic choice is the variant that
    streams H rows."""
    print("case: wide blocks choose kernel 4", flush=True)
    rng = np.random.default_rng(11)
    m, n, nnz, rank = 600, 4_000, 20_000, 32
    cells = rng.choice(m * n, size=nnz, replace=False)
    i = (cells // n).astype(np.int32)
    j = (cells % n).astype(np.int32)
    x = rng.normal(size=nnz)
    p = dict(m=m, n=n, train=(i, j, x), rank=rank, W0=rng.uniform(-0.5, 0.5, (m, rank)),
             H0=rng.uniform(-0.5, 0.5, (rank, n)))
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=2, seed=3, precision="FP32")
    o.set_arith("FP32")
    assert g.grouped_ok()[0]
    for e in range(2):
        assert g.epoch(0.01)["kernel"] == 4
        o.epoch(0.01)
        np.testing.assert_array_equal(g.get_factors()[1], o.get_factors()[1], err_msg=f"epoch {e}")
    # the batched kernel runs the same body for wide problems: bit-identical to
    # the oracle's FP32 emulation, beside a tiled problem in the same launch
    from dsgdpp_amd import batch as B
    g2, o2 = make_pair(p, update="Nzsl_L2(0.05)", tasks=2, seed=3, precision="FP32")
    small = skewed_problem(200, 160, 5000, 50, seed=5)
    s1, os1 = make_pair(small, update="Nzsl_L2(0.05)", tasks=3, seed=4, precision="FP32")
    o2.set_arith("FP32")
    os1.set_arith("FP32")
    bt = B.DsgdPpBatch([g2, s1])
    for e in range(2):
        bt.epoch([0.01, 0.02])
        o2.epoch(0.01)
        os1.epoch(0.02)
        for a, b in zip(g2.get_factors(), o2.get_factors()):
            np.testing.assert_array_equal(a, b, err_msg=f"wide, epoch {e}")
        for a, b in zip(s1.get_factors(), os1.get_factors()):
            np.testing.assert_array_equal(a, b, err_msg=f"tiled, epoch {e}")
    bt.close()


@pytest.mark.parametrize("rank,update,truncate,sgd", [(8, "Nzsl_L2(0.05)", None, "WOR"), (10, "Nzsl", None, "WR"),
                                                      (13, "Nzsl_Nzl2(0.02)", None, "SEQ"),
                                                      (37, "Nzsl_L2(0.05)", "(-0.4,0.4)", "WOR"),
                                                      (64, "Nzsl_L2(0.05)", None, "WOR"),
                                                      (100, "Nzsl_L2(0.05)", None, "WOR")])
def test_fp64_strict_grouped_equals_reference_association(rank, update, truncate, sgd):
    """FP64_STRICT on the grouped streaming kernel (the reference's per-update
    association; groups only reorder updates that share no row or column):
    bit-identical to the oracle's reference arithmetic and to the sequential
    kernel (kernel 2) on the GPU."""
    print(f"case: FP64_STRICT grouped rank {rank} {update} {truncate} {sgd}", flush=True)
    p = skewed_problem(300, 200, 12000, rank, seed=57 + rank)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=4, seed=61, sgd=sgd, precision="FP64_STRICT")
    s, _ = make_pair(p, update=update, truncate=truncate, tasks=4, seed=61, sgd=sgd, precision="FP64_STRICT",
                     kernel=2)
    ok, why = g.grouped_ok()
    assert ok, why
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == 4
        assert s.epoch(0.01)["kernel"] == 2
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        Ws, Hs = s.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Wg, Ws, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Hs, err_msg=f"epoch {e}")
        lg, lo = g.loss(0), o.loss(0)
        assert abs(lg - lo) <= 1e-12 * abs(lo), (e, lg, lo)   # the loss kernel's reduction order
    g.close()
    s.close()


@pytest.mark.parametrize("rank", [1, 13, 32, 64])
@pytest.mark.parametrize("sgd,stratum", [("WOR", "COWOR"), ("WR", "WR"), ("SEQ", "SEQ")])
def test_grouped_orders_ranks(rank, sgd, stratum):
    print(f"case: grouped rank {rank} sgd={sgd} stratum={stratum}", flush=True)
    p = skewed_problem(120, 90, 3000, rank, seed=7)
    g, o = make_pair(p, update="Nzsl_L2(0.1)", tasks=3, seed=5, precision="FP32", kernel=3,
                     sgd=sgd, stratum=stratum)
    o.set_arith("FP32")
    for e in range(2):
        g.epoch(0.02)
        o.epoch(0.02)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo)
        np.testing.assert_array_equal(Hg, Ho)


@pytest.mark.parametrize("rank,update,truncate,sgd", [(50, "Nzsl_L2(0.05)", None, "WOR"), (13, "Nzsl", None, "WR"),
                                                      (64, "Nzsl_Nzl2(0.02)", None, "WOR"),
                                                      (1, "Nzsl_L2(0.05)", "(-0.4,0.4)", "SEQ"),
                                                      (32, "Nzsl_L2(0.05)", "(-0.4,0.4)", "WOR")])
def test_two_slot_groups_bitexact_vs_emulation(monkeypatch, rank, update, truncate, sgd):
    """Every block in two-slot groups (DSGDPP_PAIRS=1): the two-slot loop keeps
    the four-slot dot-product association, so the factors equal the oracle's
    FP32 emulation bit for bit."""
    print(f"case: two-slot groups rank {rank} {update} {truncate} {sgd}", flush=True)
    monkeypatch.setenv("DSGDPP_PAIRS", "1")
    p = skewed_problem(300, 200, 12000, rank, seed=37)
    g, o = make_pair(p, update=update, truncate=truncate, tasks=4, seed=41, precision="FP32", kernel=3, sgd=sgd)
    ok, note = g.grouped_ok()
    assert ok and note.startswith("two-slot groups in ") and int(note.split()[3]) > 24, note
    o.set_arith("FP32")
    for e in range(3):
        assert g.epoch(0.01)["kernel"] == 3
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")


def test_two_slot_groups_chosen_for_chain_bound_blocks():
    """Blocks dominated by one column (here a column present in every row)
    take two-slot groups by default; the rest keep four slots; the result is
    the oracle's FP32 emulation bit for bit either way."""
    print("case: two-slot groups chosen by the chain rule", flush=True)
    rng = np.random.default_rng(3)
    m, n, r = 400, 160, 24
    dense = [(i, 5) for i in range(m)] + [(i, 97) for i in range(m)]   # two columns in every row
    other = rng.choice(m * n, size=3000, replace=False)
    cells = sorted(set(dense) | {(int(c) // n, int(c) % n) for c in other})
    i = np.array([a for a, _ in cells], dtype=np.int32)
    j = np.array([b for _, b in cells], dtype=np.int32)
    x = rng.normal(size=len(cells))
    p = dict(m=m, n=n, rank=r, train=(i, j, x), W0=rng.uniform(-0.5, 0.5, (m, r)),
             H0=rng.uniform(-0.5, 0.5, (r, n)))
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=4, seed=9, precision="FP32", kernel=3)
    ok, note = g.grouped_ok()
    assert ok and note.startswith("two-slot groups in "), note
    k = int(note.split()[3])
    assert 0 < k < 32, note
    o.set_arith("FP32")
    for e in range(3):
        g.epoch(0.01)
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")


def test_grouped_carry_over_and_many_tasks():
    print("case: grouped, equal-size blocks + d=16", flush=True)
    m, n, d = 32, 64, 4
    rng = np.random.default_rng(3)
    i, j = [], []
    for b1 in range(d):
        for b2 in range(2 * d):
            cells = rng.choice(64, size=20, replace=False)
            i += list(b1 * 8 + cells // 8)
            j += list(b2 * 8 + cells % 8)
    p = dict(m=m, n=n, train=(np.array(i), np.array(j), rng.normal(size=len(i))), rank=8,
             W0=rng.uniform(-0.5, 0.5, (m, 8)), H0=rng.uniform(-0.5, 0.5, (8, n)))
    g, o = make_pair(p, update="Nzsl_L2(0.1)", tasks=d, seed=9, precision="FP32", kernel=3)
    o.set_arith("FP32")
    for e in range(3):
        g.epoch(0.05)
        o.epoch(0.05)
        for t in range(d):
            for s_ in range(2 * d):
                np.testing.assert_array_equal(g.permutation(t, s_), o.permutation(t, s_))
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo)
        np.testing.assert_array_equal(Hg, Ho)
    p2 = skewed_problem(800, 600, 40000, 24, seed=11)
    g, o = make_pair(p2, update="Nzsl_L2(0.05)", tasks=16, seed=19, precision="FP32", kernel=3)
    o.set_arith("FP32")
    for e in range(2):
        g.epoch(0.01)
        o.epoch(0.01)
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_array_equal(Hg, Ho)


@pytest.mark.parametrize("m,n,nnz,kernel", [(120_000, 400, 300_000, 3), (20_000, 20_000, 300_000, 4)])
def test_grouped_prep_state_in_global_memory(m, n, nnz, kernel):
    """Blocks whose prep state (visit order, row and column tags) does not fit
    in LDS: the prep keeps it in global scratch, for the tile kernel (tall
    blocks, narrow column blocks) and the streaming kernel (wide column
    blocks); factors equal the oracle's FP32 emulation bit for bit."""
    print(f"case: prep state in global memory {m}x{n} nnz={nnz} -> kernel {kernel}", flush=True)
    rng = np.random.default_rng(m % 97)
    cells = np.unique(rng.integers(0, m * n, size=nnz + nnz // 10))[:nnz]
    rng.shuffle(cells)
    p = dict(m=m, n=n, train=(cells // n, cells % n, rng.normal(size=len(cells))), rank=16,
             W0=rng.uniform(-0.5, 0.5, (m, 16)), H0=rng.uniform(-0.5, 0.5, (16, n)))
    # kernel 3 asked (the grouped family: tile or streaming); on the sparse
    # square shape the automatic choice would be the level kernel
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=2, seed=17, precision="FP32", kernel=3)
    ok, why = g.grouped_ok()
    assert ok, why
    o.set_arith("FP32")
    for e in range(2):
        st = g.epoch(0.02)
        o.epoch(0.02)
        assert st["kernel"] == kernel
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
    from dsgdpp_amd import batch as B
    with pytest.raises(RuntimeError, match="prep state does not fit"):
        B.DsgdPpBatch([g])
    g.close()


@pytest.mark.parametrize("sgd", ["WOR", "WR"])
def test_grouped_prep_edge_blocks(sgd):
    """Blocks whose sizes sit on the prep's 64-point chunk boundaries (1, 2,
    63, 64, 65, 129), one block whose points all share a row (every point two
    groups after the last: the scan's window turns over many times) and one
    whose points all share a column (a chain of register-forwarded H): visit
    orders and factors equal the oracle's bit for bit."""
    print(f"case: grouped prep edge blocks sgd={sgd}", flush=True)
    R, C, d = 160, 160, 2
    m, n = d * R, 2 * d * C
    rng = np.random.default_rng(41)
    i, j = [], []
    i += [0] * 150                                  # block (0, 0): one row
    j += list(range(150))
    i += list(range(150))                           # block (0, 1): one column
    j += [C + 7] * 150
    sizes = {(0, 2): 1, (0, 3): 64, (1, 0): 63, (1, 1): 65, (1, 2): 129, (1, 3): 2}
    for (b1, b2), k in sizes.items():
        cells = rng.choice(R * C, size=k, replace=False)
        i += list(b1 * R + cells // C)
        j += list(b2 * C + cells % C)
    p = dict(m=m, n=n, train=(np.array(i), np.array(j), rng.normal(size=len(i))), rank=8,
             W0=rng.uniform(-0.5, 0.5, (m, 8)), H0=rng.uniform(-0.5, 0.5, (8, n)))
    g, o = make_pair(p, update="Nzsl_L2(0.1)", tasks=d, seed=13, precision="FP32", kernel=3, sgd=sgd,
                     stratum="COWOR")
    o.set_arith("FP32")
    for e in range(3):
        g.epoch(0.05)
        o.epoch(0.05)
        for t in range(d):
            for s_ in range(2 * d):
                np.testing.assert_array_equal(g.permutation(t, s_), o.permutation(t, s_),
                                              err_msg=f"epoch {e} task {t} sub-epoch {s_}")
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo)
        np.testing.assert_array_equal(Hg, Ho)
    g.close()


@pytest.mark.parametrize("precision,kernel", [("FP64_STRICT", 0), ("FP32", 3)])
def test_lookahead_off_matches_oracle(monkeypatch, precision, kernel):
    """Without the look-ahead plan, every epoch still draws a fresh plan in
    master-stream order (a consumed plan is never run twice)."""
    print(f"case: lookahead off {precision} kernel {kernel}", flush=True)
    monkeypatch.setenv("DSGDPP_LOOKAHEAD", "0")
    p = skewed_problem(200, 160, 5000, 16, seed=23)
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=4, seed=29, precision=precision, kernel=kernel)
    if precision == "FP32":
        o.set_arith("FP32")
    for e in range(4):
        g.epoch(0.01)
        o.epoch(0.01)
        np.testing.assert_array_equal(g.schedule(), o.schedule(), err_msg=f"epoch {e}")
        np.testing.assert_array_equal(g.task_seeds(), o.task_seeds(), err_msg=f"epoch {e}")
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")


@pytest.mark.parametrize("kind,update", [("L2", "Nzsl_L2(0.1)"), ("NZL2", "Nzsl_Nzl2(0.1)")])
def test_balance_matches_oracle(kind, update):
    """balanceSimple (balance_impl.cc:102-127) on the GPU: the device sums of
    squares feed the same multiplier as the oracle's block sums (their order
    differs, so the factors agree to 1e-14 relative, not bitwise); in the run
    loop (balance after every epoch, sgd_impl.h:110-122) FP64_STRICT losses
    stay within 1e-10 and reg(W) = reg(H) after the last epoch."""
    print(f"case: balance {kind}", flush=True)
    p = random_problem(300, 400, 9000, 8, seed=71)
    p = dict(p, W0=p["W0"] * 3.0)
    g, o = make_pair(p, update=update, tasks=4, seed=73)
    g.balance(kind)
    o.balance(kind)
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_allclose(Wg, Wo, rtol=1e-14, atol=0)
    np.testing.assert_allclose(Hg, Ho, rtol=1e-14, atol=0)
    g.close()
    g, o = make_pair(p, update=update, tasks=4, seed=73)
    rg = g.run_native(4, "BoldDriver(0.01)", balance=kind)
    ro = o.run(4, "BoldDriver(0.01)", balance={"L2": 1, "NZL2": 2}[kind])
    np.testing.assert_allclose(rg["loss"], ro["loss"], rtol=1e-10)
    np.testing.assert_array_equal(rg["eps"][1:], ro["eps"][1:])
    Wg, Hg = g.get_factors()
    if kind == "L2":
        assert (Wg ** 2).sum() == pytest.approx((Hg ** 2).sum(), rel=1e-12)
    g.close()


def test_const_decay_run_loop():
    """Const(e) (decay_constant.h:31-63) in the GPU run loop: every epoch's
    step is e, and the losses equal the oracle's run."""
    print("case: Const decay run loop", flush=True)
    p = random_problem(200, 300, 6000, 6, seed=74)
    g, o = make_pair(p, tasks=4, seed=75)
    rg = g.run_native(4, "Const(0.005)")
    ro = o.run(4, "Const(0.005)")
    assert list(rg["eps"][1:]) == [0.005] * 4
    np.testing.assert_allclose(rg["loss"], ro["loss"], rtol=LOSS_RTOL)
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_array_equal(Hg, Ho)
    g.close()


def imbalanced_problem(rank, d=4, rows=100, cols=100, heavy=3000, light=12, seed=0):
    """Row block 0's blocks hold `heavy` points each, the others `light`: the
    other tasks reach their later sub-epochs while task 0 is still in its
    first visit, so some of their dataflow waits are certainly unmet."""
    rng = np.random.default_rng(seed)
    i, j = [], []
    for b1 in range(d):
        for b2 in range(2 * d):
            cells = rng.choice(rows * cols, size=heavy if b1 == 0 else light, replace=False)
            i += list(b1 * rows + cells // cols)
            j += list(b2 * cols + cells % cols)
    m, n = d * rows, 2 * d * cols
    return dict(m=m, n=n, rank=rank, train=(np.array(i, dtype=np.int32), np.array(j, dtype=np.int32),
                                            rng.normal(size=len(i))),
                W0=rng.uniform(-0.5, 0.5, (m, rank)), H0=rng.uniform(-0.5, 0.5, (rank, n)))


@pytest.mark.parametrize("precision,kernel,arith", [("FP32", 3, "FP32"), ("BF16", 4, "BF16"),
                                                    ("FP64", 3, "FP64_GROUPED"), ("FP64_STRICT", 4, None),
                                                    ("FP64_STRICT", 2, None)])
def test_dataflow_wait_limit_reports_instead_of_hanging(monkeypatch, precision, kernel, arith):
    """Dataflow epochs are ordinary launches whose waits carry a time limit.
    Forced from the host to one tick, an unmet wait gives up: every wave of
    the launch leaves its task (tile mover and compute wave, FP64 tile,
    streaming task, sequential task), the epoch reports the error, and the
    next epoch is again bit-identical to the oracle from freshly set factors."""
    print(f"case: wait limit one tick, {precision} kernel {kernel}", flush=True)
    p = imbalanced_problem(16, seed=kernel)
    g, o = make_pair(p, update="Nzsl_L2(0.05)", tasks=4, seed=5, precision=precision, kernel=kernel)
    if arith:
        o.set_arith(arith)
    monkeypatch.setenv("DSGDPP_WAIT_LIMIT_MS", "0.00001")
    with pytest.raises(RuntimeError, match="dataflow wait gave up"):
        g.epoch(0.01)
    monkeypatch.delenv("DSGDPP_WAIT_LIMIT_MS")
    o.epoch(0.01)                                 # the oracle's streams pass the failed epoch too
    g.set_factors(p["W0"], p["H0"])
    o.set_factors(p["W0"], p["H0"])
    if arith in ("FP32", "BF16"):
        o.set_arith(arith)                        # factors as the GPU stores them
    for e in range(2):
        assert g.epoch(0.01)["kernel"] == kernel
        o.epoch(0.01)
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
    g.close()


@pytest.mark.parametrize("mode", ["1", "0"])
def test_grouped_prep_visit_orders_either_way(monkeypatch, mode):
    """The WOR visit orders of the grouped prep come from the parallel
    visit-order kernel (DSGDPP_PREP_ORDER=1, the batches' default) or from the
    prep itself (0, the single contexts' default): the orders equal the
    oracle's and the factors its FP32 emulation, bit for bit, either way,
    including orders that carry over between equal-size blocks."""
    print(f"case: grouped prep, DSGDPP_PREP_ORDER={mode}", flush=True)
    monkeypatch.setenv("DSGDPP_PREP_ORDER", mode)
    p = skewed_problem(400, 300, 9000, 24, seed=77)
    g, o = make_pair(p, update="Nzsl_L2(0.1)", tasks=3, seed=61, precision="FP32", kernel=3)
    o.set_arith("FP32")
    for e in range(3):
        assert g.epoch(0.02)["kernel"] in (3, 4)
        o.epoch(0.02)
        for t in range(3):
            for s in (0, 2, 5):
                np.testing.assert_array_equal(g.permutation(t, s), o.permutation(t, s), err_msg=f"{e} {t} {s}")
        np.testing.assert_array_equal(g.get_factors()[0], o.get_factors()[0], err_msg=f"epoch {e}")
        np.testing.assert_array_equal(g.get_factors()[1], o.get_factors()[1], err_msg=f"epoch {e}")
    g.close()
    # equal block sizes: the orders carry over from block to block
    rng = np.random.default_rng(5)
    d, rows, cols = 2, 30, 30
    i, j = [], []
    for b1 in range(d):
        for b2 in range(2 * d):
            cells = rng.choice(rows * cols, size=200, replace=False)
            i += list(b1 * rows + cells // cols)
            j += list(b2 * cols + cells % cols)
    q = dict(m=d * rows, n=2 * d * cols, train=(np.array(i), np.array(j), rng.normal(size=len(i))), rank=8,
             W0=rng.uniform(-0.5, 0.5, (d * rows, 8)), H0=rng.uniform(-0.5, 0.5, (8, 2 * d * cols)))
    g, o = make_pair(q, update="Nzsl_L2(0.1)", tasks=d, seed=3, precision="FP32", kernel=3)
    o.set_arith("FP32")
    for e in range(2):
        g.epoch(0.05)
        o.epoch(0.05)
        for t in range(d):
            for s in range(2 * d):
                np.testing.assert_array_equal(g.permutation(t, s), o.permutation(t, s))
        np.testing.assert_array_equal(g.get_factors()[1], o.get_factors()[1], err_msg=f"carry-over epoch {e}")
    g.close()


@pytest.mark.parametrize("scan", ["seq", "rounds"])
@pytest.mark.parametrize("precision,kernel,pairs,lookahead", [("FP32", 3, "0", "1"), ("FP32", 3, "1", "0"),
                                                               ("FP32", 4, "0", "1"), ("BF16", 3, "0", "0"),
                                                               ("FP64", 3, "0", "1")])
def test_grouped_prep_scans_either_way(monkeypatch, scan, precision, kernel, pairs, lookahead):
    """The grouped prep places the points in groups with the sequential scan
    (a look-ahead prep's default) or the round scan (the points of a chunk in
    parallel; a prep that the epoch waits for): any grouping that keeps each
    row's and column's visit order gives the same factors, so both equal the
    oracle's emulation bit for bit.  Covered: the LDS tile and the streaming
    kernel (H forwarding entries), two-slot visits, look-ahead on and off,
    FP32, BF16 and FP64, on cells with hot rows and columns."""
    print(f"case: grouped prep scan={scan} {precision} kernel {kernel} pairs {pairs} lookahead {lookahead}",
          flush=True)
    monkeypatch.setenv("DSGDPP_PREP_SCAN", scan)
    monkeypatch.setenv("DSGDPP_PAIRS", pairs)
    monkeypatch.setenv("DSGDPP_LOOKAHEAD", lookahead)
    p = skewed_problem(600, 400, 14000, 16, seed=91)
    g, o = make_pair(p, update="Nzsl_L2(0.1)", tasks=3, seed=17, precision=precision, kernel=kernel)
    o.set_arith({"FP32": "FP32", "BF16": "BF16", "FP64": "FP64_GROUPED"}[precision])
    for e in range(3):
        assert g.epoch(0.02)["kernel"] == kernel
        o.epoch(0.02)
        np.testing.assert_array_equal(g.get_factors()[0], o.get_factors()[0], err_msg=f"W epoch {e}")
        np.testing.assert_array_equal(g.get_factors()[1], o.get_factors()[1], err_msg=f"H epoch {e}")
    g.close()


@pytest.mark.parametrize("scan", ["seq", "rounds"])
def test_grouped_prep_scans_32bit_tags(monkeypatch, scan):
    """Blocks of 8,000 points and more: the prep keeps its tags and visit
    orders in 32 bits.  Both scans give factors bit-identical to the oracle's
    FP32 emulation, look-ahead on and off."""
    print(f"case: grouped prep scan={scan}, 20,000-point blocks (32-bit tags)", flush=True)
    monkeypatch.setenv("DSGDPP_PREP_SCAN", scan)
    p = skewed_problem(600, 400, 40000, 16, seed=13)
    for la in ("1", "0"):
        monkeypatch.setenv("DSGDPP_LOOKAHEAD", la)
        g, o = make_pair(p, update="Nzsl_L2(0.1)", tasks=1, seed=23, precision="FP32", kernel=3)
        o.set_arith("FP32")
        assert g.block_nnz().max() >= 8000
        for e in range(2):
            assert g.epoch(0.01)["kernel"] == 3
            o.epoch(0.01)
            np.testing.assert_array_equal(g.get_factors()[0], o.get_factors()[0], err_msg=f"W epoch {e} la {la}")
            np.testing.assert_array_equal(g.get_factors()[1], o.get_factors()[1], err_msg=f"H epoch {e} la {la}")
        g.close()
