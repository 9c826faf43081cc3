"""The level-parallel epoch (kernel 5) against the oracle, through the C ABI.

A block runs level by level (a point's level is one more than the levels of
the previous points on its row and on its column in visit order), one lane
per point.  Every W row and H column then sees the sequential update
sequence, so the factors are bit-identical to the oracle in each precision's
own association: FP64_STRICT to the reference arithmetic, FP64 to the FP64
grouped emulation, FP32/BF16 to the float emulations.  Covered: the three
update functors with and without truncation, WOR and SEQ point orders,
permutation carry-over, empty and one-point blocks, a block with a column of
9,000 points (more levels than one pass of the prep's level histogram), the
automatic choice on the reference's quick-start shape, and the time limit of
the dataflow waits.
"""
import numpy as np
import pytest

from conftest import random_problem

from test_gpu_parity import imbalanced_problem, make_pair

pytestmark = pytest.mark.gpu

ARITH = {"FP64_STRICT": None, "FP64": "FP64_GROUPED", "FP32": "FP32", "BF16": "BF16"}


def _check(g, o, e):
    Wg, Hg = g.get_factors()
    Wo, Ho = o.get_factors()
    np.testing.assert_array_equal(Wg, Wo, err_msg=f"W epoch {e}")
    np.testing.assert_array_equal(Hg, Ho, err_msg=f"H epoch {e}")
    lg, lo = g.loss(0), o.loss(0)
    assert abs(lg - lo) <= 1e-10 * abs(lo), (e, lg, lo)


def _pair(p, precision, **kw):
    g, o = make_pair(p, precision=precision, kernel=5, **kw)
    if ARITH[precision]:
        o.set_arith(ARITH[precision])
    return g, o


@pytest.mark.parametrize("precision", ["FP64_STRICT", "FP64", "FP32", "BF16"])
@pytest.mark.parametrize("update,truncate", [("Nzsl", None), ("Nzsl_L2(0.5)", None), ("Nzsl_Nzl2(0.05)", None),
                                             ("Nzsl_L2(0.5)", "(-0.3,0.3)")])
def test_level_bitexact(precision, update, truncate):
    print(f"case: level kernel {precision} {update} truncate={truncate}", flush=True)
    p = random_problem(3000, 2400, 60_000, 10, seed=31)
    g, o = _pair(p, precision, update=update, truncate=truncate, tasks=2, seed=19)
    for e, eps in enumerate((0.02, 0.03)):
        assert g.epoch(eps)["kernel"] == 5
        o.epoch(eps)
        _check(g, o, e)
    g.close()


@pytest.mark.parametrize("precision,rank", [("FP64_STRICT", 1), ("FP64", 16), ("FP32", 17), ("FP32", 32),
                                            ("BF16", 24)])
def test_level_ranks(precision, rank):
    print(f"case: level kernel rank {rank} {precision}", flush=True)
    p = random_problem(1500, 1200, 30_000, rank, seed=rank)
    g, o = _pair(p, precision, update="Nzsl_L2(0.2)", tasks=3, seed=rank + 3)
    for e in range(2):
        assert g.epoch(0.02)["kernel"] == 5
        o.epoch(0.02)
        _check(g, o, e)
    g.close()


@pytest.mark.parametrize("sgd,stratum", [("SEQ", "COWOR"), ("WOR", "SEQ"), ("WOR", "WR")])
def test_level_orders(sgd, stratum):
    print(f"case: level kernel sgd={sgd} stratum={stratum}", flush=True)
    p = random_problem(2000, 2000, 40_000, 8, seed=5)
    g, o = _pair(p, "FP64_STRICT", update="Nzsl_L2(1)", tasks=4, seed=29, sgd=sgd, stratum=stratum)
    for e in range(2):
        assert g.epoch(0.01)["kernel"] == 5
        o.epoch(0.01)
        _check(g, o, e)
        if sgd == "WOR":
            for t in range(4):
                np.testing.assert_array_equal(g.permutation(t, 3), o.permutation(t, 3))
    g.close()


def test_level_carry_over_empty_and_single_point_blocks():
    """Equal block sizes (each task's permutation carries over), empty blocks
    and blocks of one point."""
    print("case: level kernel carry-over, empty and one-point blocks", flush=True)
    rng = np.random.default_rng(3)
    d, rows, cols = 3, 40, 20
    i, j = [], []
    for b1 in range(d):
        for b2 in range(2 * d):
            k = 0 if (b1 + b2) % 5 == 0 else (1 if (b1 + b2) % 5 == 1 else 300)
            cells = rng.choice(rows * cols, size=k, replace=False)
            i += list(b1 * rows + cells // cols)
            j += list(b2 * cols + cells % cols)
    m, n = d * rows, 2 * d * cols
    p = dict(m=m, n=n, train=(np.array(i), np.array(j), rng.normal(size=len(i))), rank=6,
             W0=rng.uniform(-0.5, 0.5, (m, 6)), H0=rng.uniform(-0.5, 0.5, (6, n)))
    for precision in ("FP64_STRICT", "FP32"):
        g, o = _pair(p, precision, update="Nzsl_L2(0.1)", tasks=d, seed=37)
        nz = g.block_nnz()
        assert (nz == 0).any() and (nz == 1).any() and (nz == 300).sum() >= 6
        for e in range(3):
            assert g.epoch(0.05)["kernel"] == 5
            o.epoch(0.05)
            _check(g, o, e)
        g.close()


def test_level_deep_column_many_histogram_passes():
    """One column block holds a column of about 9,000 points: its visits have
    more levels than one pass of the prep's level histogram covers."""
    print("case: level kernel, a 9,000-point column (deep levels)", flush=True)
    rng = np.random.default_rng(11)
    m, n = 20_000, 8
    rows = rng.choice(m, size=9_000, replace=False)
    i = list(rows)
    j = [0] * rows.size
    cells = rng.choice(m * n, size=6_000, replace=False)
    keep = (cells % n) != 0
    i += list(cells[keep] // n)
    j += list(cells[keep] % n)
    p = dict(m=m, n=n, train=(np.array(i), np.array(j), rng.normal(size=len(i))), rank=4,
             W0=rng.uniform(-0.5, 0.5, (m, 4)), H0=rng.uniform(-0.5, 0.5, (4, n)))
    g, o = _pair(p, "FP64_STRICT", update="Nzsl_L2(0.1)", tasks=1, seed=3)
    for e in range(2):
        assert g.epoch(0.01)["kernel"] == 5
        o.epoch(0.01)
        _check(g, o, e)
    g.close()


def test_level_unsupported_cases_are_refused():
    import dsgdpp_amd as D
    p = random_problem(500, 500, 5000, 20, seed=2)
    g = D.DsgdPp(update="Nzsl_L2(1)", loss="Nzsl_L2(1)", rank=20, tasks_per_rank=2, precision="FP64", kernel=5)
    with pytest.raises(RuntimeError, match="level kernel"):
        g.load_coo(D.TRAIN, p["m"], p["n"], *p["train"])
    g.close()
    g = D.DsgdPp(update="Nzsl_L2(1)", loss="Nzsl_L2(1)", rank=4, tasks_per_rank=2, sgd_order="WR", kernel=5)
    with pytest.raises(RuntimeError, match="level kernel"):
        g.load_coo(D.TRAIN, p["m"], p["n"], *p["train"])
    g.close()


def test_level_auto_choice():
    """Kernel 0 takes the level kernel on large blocks whose rows and columns
    hold few points each (the quick-start shape at 1/5 of each dimension), and
    keeps the grouped kernel on small or skewed blocks."""
    from dsgdpp_amd import data
    print("case: automatic choice of the level kernel", flush=True)
    p = data.quickstart_1m(seed=3, shrink=5)
    g, o = make_pair(p, update="Nzsl_L2(1)", tasks=4, seed=41, precision="FP64_STRICT")
    for e in range(2):
        assert g.epoch(0.001)["kernel"] == 5
        o.epoch(0.001)
        _check(g, o, e)
    g.close()
    small = random_problem(300, 200, 3000, 8, seed=4)
    g, _ = make_pair(small, update="Nzsl_L2(1)", tasks=2, seed=1, precision="FP32")
    assert g.epoch(0.01)["kernel"] == 3
    g.close()


def test_level_wait_limit_reports_instead_of_hanging(monkeypatch):
    """Forced from the host to one tick, an unmet dataflow wait of the level
    kernel gives up: every workgroup leaves, the epoch reports the error, and
    the next epochs are again bit-identical from freshly set factors."""
    print("case: level kernel, wait limit one tick", flush=True)
    p = imbalanced_problem(8, seed=9)
    g, o = _pair(p, "FP64_STRICT", update="Nzsl_L2(0.05)", tasks=4, seed=5)
    monkeypatch.setenv("DSGDPP_WAIT_LIMIT_MS", "0.00001")
    with pytest.raises(RuntimeError, match="dataflow wait gave up"):
        g.epoch(0.01)
    monkeypatch.delenv("DSGDPP_WAIT_LIMIT_MS")
    o.epoch(0.01)
    g.set_factors(p["W0"], p["H0"])
    o.set_factors(p["W0"], p["H0"])
    for e in range(2):
        assert g.epoch(0.01)["kernel"] == 5
        o.epoch(0.01)
        _check(g, o, e)
    g.close()
