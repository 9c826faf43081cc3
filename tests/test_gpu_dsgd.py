"""DSGD (mf/sgd/dsgd_impl.h:81-234, 329-350; tools/mfdsgd.cc) on the GPU
against the oracle's direct restatement of it (d x d blocks, d sub-epochs).

The library runs DSGD on its DSGD++ kernels: DSGD's column block k is block 2k
of a d x 2d grid whose odd blocks are empty, and DSGD's sub-epoch s is
sub-epoch 2s.  These tests check that this mapping is exact: d x d schedules,
task seeds and every visit order equal the oracle's, FP64_STRICT factors are
bitwise equal (and FP32/BF16/FP64 equal the oracle's emulation of the grouped
kernels), on every kernel, every point and stratum order, every functor, in
the run loop, the mfdsgd command line and a batch.
"""
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import random_problem
from oracle import pyoracle as po
from test_gpu_parity import LOSS_RTOL, skewed_problem

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def dsgd_pair(p, update="Nzsl_L2(1)", truncate=None, tasks=3, world=1, sgd="WOR", stratum="COWOR", seed=7,
              precision="FP64_STRICT", kernel=0, test=None, threads=2):
    import dsgdpp_amd as D
    g = D.DsgdPp(update=update, loss=update, rank=p["rank"], tasks_per_rank=tasks, world_size=world,
                 sgd_order=sgd, stratum_order=stratum, seed=seed, precision=precision, truncate=truncate,
                 kernel=kernel, algorithm="DSGD")
    o = po.Oracle(update=update, loss=update, truncate=truncate, rank=p["rank"], world_size=world,
                  tasks_per_rank=tasks, sgd_order=sgd, stratum_order=stratum, seed=seed, threads=threads,
                  algorithm="DSGD")
    for h in (g, o):
        h.load_coo(0, p["m"], p["n"], *p["train"])
        if test is not None:
            h.load_coo(1, p["m"], p["n"], *test)
        h.set_factors(p["W0"], p["H0"])
    o.record(True)
    return g, o


def assert_dsgd_epoch(g, o, eps, perms=True):
    g.epoch(eps)
    o.epoch(eps)
    d = g.d
    assert g.schedule().shape == (d, d)
    np.testing.assert_array_equal(g.schedule(), o.schedule())
    np.testing.assert_array_equal(g.task_seeds(), o.task_seeds())
    if perms:
        for t in range(d):
            for s in range(d):
                np.testing.assert_array_equal(g.permutation(t, s), o.permutation(t, s),
                                              err_msg=f"task {t} sub-epoch {s}")
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_array_equal(Hg, Ho)


@pytest.mark.parametrize("sgd,stratum", list(itertools.product(["SEQ", "WR", "WOR"],
                                                               ["SEQ", "RSEQ", "WR", "WOR", "COWOR"])))
def test_dsgd_orders_strict(sgd, stratum):
    print(f"case: DSGD orders sgd={sgd} stratum={stratum}", flush=True)
    p = random_problem(40, 60, 500, 4, seed=2)
    g, o = dsgd_pair(p, tasks=3, sgd=sgd, stratum=stratum, seed=11)
    np.testing.assert_array_equal(g.block_nnz(), o.block_nnz())
    for e in range(3):
        assert_dsgd_epoch(g, o, 0.02)
    lg, lo = g.loss(0), o.loss(0)
    assert abs(lg - lo) <= LOSS_RTOL * abs(lo)
    g.close()


@pytest.mark.parametrize("update,truncate", [("Nzsl", None), ("Nzsl_L2(0.5)", None), ("Nzsl_Nzl2(0.05)", None),
                                             ("Nzsl_L2(0.5)", "(-0.3,0.3)")])
@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_dsgd_functors_kernels_strict(update, truncate, kernel):
    print(f"case: DSGD functor {update} truncate={truncate} kernel {kernel}", flush=True)
    p = random_problem(50, 70, 900, 5, seed=3)
    g, o = dsgd_pair(p, update=update, truncate=truncate, tasks=4, seed=5, kernel=kernel)
    for e in range(3):
        assert_dsgd_epoch(g, o, 0.05, perms=(e == 0))
    g.close()


def test_dsgd_cowor_world_groups_and_empty_blocks_strict():
    print("case: DSGD COWOR w=2 t=3, empty blocks and duplicates", flush=True)
    p = random_problem(60, 80, 1200, 4, seed=4)
    g, o = dsgd_pair(p, tasks=3, world=2, seed=9)
    for e in range(2):
        assert_dsgd_epoch(g, o, 0.03)
    g.close()
    q = random_problem(64, 64, 40, 3, seed=6, dup=10)   # d = 8: 64 blocks, most empty
    g, o = dsgd_pair(q, tasks=8, seed=13, sgd="WR")
    np.testing.assert_array_equal(g.block_nnz(), o.block_nnz())
    for e in range(3):
        assert_dsgd_epoch(g, o, 0.05)
    g.close()


@pytest.mark.parametrize("precision,arith,kernel", [("FP32", "FP32", 3), ("BF16", "BF16", 3), ("FP32", "FP32", 4),
                                                    ("FP64", "FP64_GROUPED", 3), ("FP64_STRICT", None, 4)])
def test_dsgd_grouped_kernels(precision, arith, kernel):
    """The grouped kernels on DSGD's blocking, against the oracle's emulation
    of their association (FP64_STRICT: the reference association)."""
    print(f"case: DSGD grouped {precision} kernel {kernel}", flush=True)
    p = skewed_problem(300, 200, 12000, 37, seed=91)
    g, o = dsgd_pair(p, update="Nzsl_L2(0.05)", tasks=4, seed=43, precision=precision, kernel=kernel)
    if arith:
        o.set_arith(arith)
    for e in range(3):
        st = g.epoch(0.01)
        assert st["kernel"] == kernel
        o.epoch(0.01)
        np.testing.assert_array_equal(g.permutation(2, 3), o.permutation(2, 3))
        Wg, Hg = g.get_factors()
        Wo, Ho = o.get_factors()
        np.testing.assert_array_equal(Wg, Wo, err_msg=f"epoch {e}")
        np.testing.assert_array_equal(Hg, Ho, err_msg=f"epoch {e}")
    g.close()


def test_dsgd_quickstart_shape_run_loop():
    """The reference's quick-start shape (reduced 4x) as mfdsgd would run it
    with --tasks-per-rank 4 (4 x 4 blocks): the defaultRunner loop, 8 epochs,
    per-epoch training and test loss within 1e-9, factors bitwise equal."""
    print("case: DSGD quick-start shape / 4, run loop, 8 epochs", flush=True)
    from dsgdpp_amd import data
    p = data.quickstart_1m(shrink=4, seed=3)
    g, o = dsgd_pair(p, update="Nzsl_L2(1)", tasks=4, seed=1234, test=p["test"], threads=4)
    rg = g.run_native(8, "BoldDriver(0.001)")
    ro = o.run(8, "BoldDriver(0.001)")
    np.testing.assert_allclose(rg["loss"], ro["loss"], rtol=1e-9)
    np.testing.assert_allclose(rg["test_loss"], ro["test_loss"], rtol=1e-9)
    np.testing.assert_array_equal(rg["eps"][1:], ro["eps"][1:])
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo)
    np.testing.assert_array_equal(Hg, Ho)
    assert ro["loss"][-1] < ro["loss"][0]
    g.close()


def test_dsgd_python_runner_mirror():
    """DsgdJob / DsgdRunner (mf/sgd/dsgd.h) drive the same context kind."""
    print("case: DSGD Python runner mirror", flush=True)
    import dsgdpp_amd as D
    p = random_problem(120, 90, 2000, 6, seed=15)
    job = D.DsgdJob(p["m"], p["n"], p["train"], p["W0"], p["H0"], D.UpdateNzslL2(0.1), tasks_per_rank=3, seed=21)
    tr = D.DsgdRunner().run(job, 3, D.BoldDriver(0.01))
    o = po.Oracle(update="Nzsl_L2(0.1)", loss="Nzsl_L2(0.1)", rank=6, tasks_per_rank=3, seed=21, algorithm="DSGD")
    o.load_coo(0, p["m"], p["n"], *p["train"])
    o.set_factors(p["W0"], p["H0"])
    ro = o.run(3, "BoldDriver(0.01)")
    np.testing.assert_allclose([e.loss for e in tr.entries], ro["loss"], rtol=LOSS_RTOL)
    with pytest.raises(TypeError):
        D.DsgdRunner().epoch(D.DsgdPpJob(p["m"], p["n"], p["train"], p["W0"], p["H0"], D.UpdateNzslL2(0.1)), 0.01)
    job.ctx.close()


def test_dsgd_batch_vs_oracle():
    """DSGD contexts batched with a DSGD++ context in one launch."""
    print("case: DSGD contexts in a batch", flush=True)
    import dsgdpp_amd as D
    from dsgdpp_amd import batch as B
    ctxs, orcs = [], []
    for k, (algo, r, d) in enumerate([("DSGD", 16, 3), ("DSGDPP", 32, 2), ("DSGD", 64, 4)]):
        p = skewed_problem(300, 240, 6000, r, seed=60 + k)
        g = D.DsgdPp(update="Nzsl_L2(0.05)", loss="Nzsl_L2(0.05)", rank=r, tasks_per_rank=d, seed=k,
                     precision="FP32", algorithm=algo)
        o = po.Oracle(update="Nzsl_L2(0.05)", loss="Nzsl_L2(0.05)", rank=r, tasks_per_rank=d, seed=k,
                      algorithm=algo)
        for h in (g, o):
            h.load_coo(0, p["m"], p["n"], *p["train"])
            h.set_factors(p["W0"], p["H0"])
        o.set_arith("FP32")
        ctxs.append(g)
        orcs.append(o)
    bt = B.DsgdPpBatch(ctxs)
    for e in range(3):
        bt.epoch([0.01, 0.02, 0.015])
        for k, (g, o) in enumerate(zip(ctxs, orcs)):
            o.epoch([0.01, 0.02, 0.015][k])
            for a, b in zip(g.get_factors(), o.get_factors()):
                np.testing.assert_array_equal(a, b, err_msg=f"problem {k} epoch {e}")
    bt.close()
    for g in ctxs:
        g.close()


def test_mfdsgd_cli_matches_oracle_cli(tmp_path):
    """bin/mfdsgd (the program name selects DSGD) against the oracle CLI with
    --algorithm DSGD: byte-identical factor files and the same trace losses;
    --map-reduce is accepted (one GPU: the same updates)."""
    print("case: mfdsgd CLI vs oracle CLI", flush=True)
    from test_gpu_files import _losses, _write
    _write(tmp_path, 60, 80, 4, 2000, 300, seed=7)
    common = ["--input-file", str(tmp_path / "v.mmc"), "--input-test-file", str(tmp_path / "t.mmc"),
              "--input-row-file", str(tmp_path / "w.mma"), "--input-col-file", str(tmp_path / "h.mma"),
              "--update", "Nzsl_L2(0.1)", "--regularize", "None", "--loss", "Nzsl_L2(0.1)", "--rank", "4",
              "--decay", "BoldDriver(0.05)", "--epochs", "4", "--tasks-per-rank", "3", "--seed", "19"]
    runs = {"gpu": [os.path.join(ROOT, "dsgdpp_amd", "bin", "mfdsgd"), "--map-reduce", "false"],
            "cpu": [os.path.join(ROOT, "oracle", "mfdsgdpp_cpu"), "--algorithm", "DSGD"]}
    for tag, cmd in runs.items():
        r = subprocess.run(cmd + common + ["--output-row-file", str(tmp_path / f"w_{tag}.mma"),
                                           "--output-col-file", str(tmp_path / f"h_{tag}.mma"),
                                           "--trace", str(tmp_path / f"t_{tag}.R")],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (tag, r.stderr)
        if tag == "gpu":
            assert r.stdout.startswith("DSGD options")
    for f in ("w", "h"):
        assert (tmp_path / f"{f}_gpu.mma").read_text() == (tmp_path / f"{f}_cpu.mma").read_text(), f
    lg = _losses((tmp_path / "t_gpu.R").read_text())
    assert len(lg) == 5 and lg == _losses((tmp_path / "t_cpu.R").read_text())
