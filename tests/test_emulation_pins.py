"""Pins of the oracle's emulations of the GPU's grouped update -- FP32 factors
and math, BF16 factor storage with FP32 math, FP64 with fused multiply-adds --
against the mathematics of the reference's update functors (UpdateNzsl,
update-nzsl.h:34-49; UpdateNzslL2, update-nzsl-l2.h:45-55; UpdateNzslNzl2,
update-nzsl-nzl2.h:34-53; UpdateTruncate, update-truncate.h:46-66).

The GPU kernels equal these emulations bit for bit (tests/test_gpu_parity.py),
so what pins the emulations pins the kernels.

Per update, every element of an emulated result lies within a bound derived
from the emulation's own rounding steps, under the standard model
fl(a op b) = (a op b)(1 + d), |d| <= u:
  - u = 2^-24 for FP32 math;
  - for BF16 storage, plus 2^-8 relative for the final rounding to bf16;
  - u = 2^-53 for the FP64 emulation.
The exact value is evaluated in rational arithmetic from the same inputs.

Per epoch, the emulations' training losses stay within 1e-6 (FP32), 2e-2
(BF16) and 1e-12 (FP64) relative of the reference arithmetic's, on the same
schedules and visit orders.

Each check is shown to fail on planted mistakes in the emulation: the
regularization term dropped, h updated with the new w, and the row and column
counts swapped.  BF16's per-epoch tolerance is too loose to see them; its
per-update bound sees all three.
"""
from fractions import Fraction

import numpy as np
import pytest

from conftest import random_problem
from oracle import pyoracle as po

U = {"FP32": 2.0 ** -24, "BF16": 2.0 ** -24, "FP64_GROUPED": 2.0 ** -53}
U_BF16 = 2.0 ** -8
KINDS = ("Nzsl", "Nzsl_L2", "Nzsl_Nzl2")


def _bf16(v):
    """Round-to-nearest-even to bfloat16 (8 significant bits), as numpy float64."""
    b = np.asarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _cases(arith, seed=0, count=144):
    """Update inputs as the GPU stores them: factors representable in the
    storage type; x, eps and lambda as doubles (the FP32/BF16 kernels round
    them to float themselves)."""
    rng = np.random.default_rng(seed)
    ranks = [1, 5, 16, 50, 64, 65, 100, 128]
    out = []
    for k in range(count):
        r = ranks[k % len(ranks)]
        kind = KINDS[(k // len(ranks)) % 3]
        trunc = (k // (3 * len(ranks))) % 2 == 1
        s = [0.05, 0.3, 1.0, 3.0][k % 4]
        w = rng.normal(0, s, r)
        h = rng.normal(0, s, r)
        if arith == "FP32":
            w, h = w.astype(np.float32).astype(np.float64), h.astype(np.float32).astype(np.float64)
        elif arith == "BF16":
            w, h = _bf16(w), _bf16(h)
        x = float(rng.integers(1, 6)) if k % 3 == 0 else float(rng.normal(0, 3))
        eps = float(10 ** rng.uniform(-3, -0.5))
        lam = float(10 ** rng.uniform(-2, 0.3))
        n1, n2 = float(rng.integers(1, 60)), float(rng.integers(1, 60))
        tmin, tmax = (-0.4 * s, 0.4 * s) if trunc else (0.0, 0.0)
        out.append(dict(kind=kind, lam=lam, w=w, h=h, x=x, eps=eps, n1=n1, n2=n2, tmin=tmin, tmax=tmax,
                        truncate=trunc))
    return out


def _exact(c):
    """The update in exact rational arithmetic, before truncation: wh = sum
    w_z h_z; f1 = -2 eps (x - wh); w'_z = w_z - f1 h_z - a w_z and
    h'_z = h_z - f1 w_z - b h_z with the old w_z, where a = 2 eps lambda / nnz1,
    b = 2 eps lambda / nnz2 (Nzsl_L2), a = b = 2 eps lambda (Nzsl_Nzl2),
    a = b = 0 (Nzsl)."""
    F = Fraction
    w = [F(v) for v in c["w"]]
    h = [F(v) for v in c["h"]]
    wh = sum((a * b for a, b in zip(w, h)), F(0))
    f1 = -2 * F(c["eps"]) * (F(c["x"]) - wh)
    f2 = 2 * F(c["eps"]) * F(c["lam"])
    if c["kind"] == "Nzsl_L2":
        a, b = f2 / F(c["n1"]), f2 / F(c["n2"])
    elif c["kind"] == "Nzsl_Nzl2":
        a = b = f2
    else:
        a = b = F(0)
    wn = [wz - f1 * hz - a * wz for wz, hz in zip(w, h)]
    hn = [hz - f1 * wz - b * hz for wz, hz in zip(w, h)]
    return (np.array([float(v) for v in wn]), np.array([float(v) for v in hn]),
            abs(float(f1)), float(a), float(b), float(wh))


def _bound(c, arith, f1, a, b, wh):
    """Elementwise error bounds of the emulated w' and h' (before truncation
    and the BF16 store), from the emulation's rounding steps:
      dot product: lane chains of E terms (E = 4 up to rank 64, else 8), then
        a 4-level pairwise tree, so at most E + 4 roundings on any path:
        |err| <= gamma_(E+4) sum |w_z h_z|;
      x - wh: x rounded to the working precision (FP32, BF16), one rounding;
      f1 = (eps * -2) * (x - wh): eps rounded (FP32, BF16), one product;
      a, b: eps, lambda, the product, 1/nnz (in double, then to float) and the
        last product: at most 6 u relative;
      w' = fma(-f1, h, fma(-a, w, w)): two roundings, of |w - a w| and of
        |w'| <= |w| + a |w| + |f1 h|."""
    u = U[arith]
    w, h, x, eps = np.abs(c["w"]), np.abs(c["h"]), abs(c["x"]), c["eps"]
    E = 8 if len(w) > 64 else 4
    k = E + 4
    S = float(np.sum(w * h))
    dwh = k * u / (1 - k * u) * S
    ux = 0.0 if arith == "FP64_GROUPED" else u * x
    D = abs(c["x"] - wh)
    ed = ux + dwh + u * (D + ux + dwh)
    df1 = 2 * eps * (ed * (1 + 3 * u) + D * 2.1 * u)
    da, db = 6 * u * a, 6 * u * b
    bw = w * da + h * df1 + 2 * u * (w + a * w + f1 * h)
    bh = h * db + w * df1 + 2 * u * (h + b * h + f1 * w)
    return 1.25 * bw + 1e-300, 1.25 * bh + 1e-300


def _violations(c, arith, got_w, got_h):
    """Elements of the emulated result outside the bound (0 = all inside)."""
    ew, eh, f1, a, b, wh = _exact(c)
    bw, bh = _bound(c, arith, f1, a, b, wh)
    u = U[arith]
    bad = 0
    for exact, bnd, got in ((ew, bw, got_w), (eh, bh, got_h)):
        cands = [(exact, bnd)]
        if c["truncate"]:
            lo, hi, eps = c["tmin"], c["tmax"], c["eps"]
            tb = 2.5 * u * (abs(lo) + abs(hi) + eps) + 1e-300
            # thresholds compared in the working precision: within the bound
            # (plus the rounding of tmin/tmax) either branch may be taken
            near_lo = np.abs(exact - lo) <= bnd + u * abs(lo)
            near_hi = np.abs(exact - hi) <= bnd + u * abs(hi)
            t = np.where(exact <= lo, lo + eps, np.where(exact >= hi, hi - eps, exact))
            tb_ = np.where((exact <= lo) | (exact >= hi), tb, bnd)
            alt_lo = np.where(near_lo, lo + eps, t)
            alt_hi = np.where(near_hi, hi - eps, t)
            alt_id = np.where(near_lo | near_hi, exact, t)
            cands = [(t, tb_), (alt_lo, np.maximum(tb_, tb)), (alt_hi, np.maximum(tb_, tb)),
                     (alt_id, np.maximum(tb_, bnd))]
        ok = np.zeros(len(got), dtype=bool)
        for val, bd in cands:
            if arith == "BF16":   # the stored value is the float result rounded to bf16
                bd = bd + U_BF16 * (np.abs(val) + bd)
            ok |= np.abs(got - val) <= bd
        bad += int((~ok).sum())
    return bad


def _run(c, arith, mutant=0):
    return po.update_point_arith(arith, c["kind"], c["lam"], c["w"], c["h"], c["x"], c["eps"], c["n1"], c["n2"],
                                 c["tmin"], c["tmax"], c["truncate"], mutant=mutant)


def test_reference_update_matches_exact_value():
    """The oracle's FP64 reference update (pinned by known answers and an
    exact-rational check of every rounding in tests/test_oracle.py) lies
    within its own rounding of the exact rational value, so exact value and
    reference describe the same update."""
    for c in _cases("FP64_GROUPED", seed=5, count=48):
        c = dict(c, truncate=False)
        w, h = _run(c, "FP64")
        ew, eh, f1, a, b, wh = _exact(c)
        r = len(c["w"])
        # sequential dot over r terms, then a handful of roundings per element
        g = (r + 8) * 2.0 ** -53
        S = float(np.sum(np.abs(c["w"] * c["h"])))
        tol_w = 8 * 2.0 ** -53 * (np.abs(c["w"]) * (1 + a) + f1 * np.abs(c["h"])) + \
            2 * c["eps"] * g * (S + abs(c["x"])) * np.abs(c["h"]) * 2 + 1e-300
        tol_h = 8 * 2.0 ** -53 * (np.abs(c["h"]) * (1 + b) + f1 * np.abs(c["w"])) + \
            2 * c["eps"] * g * (S + abs(c["x"])) * np.abs(c["w"]) * 2 + 1e-300
        assert (np.abs(w - ew) <= tol_w).all()
        assert (np.abs(h - eh) <= tol_h).all()


@pytest.mark.parametrize("arith", ["FP32", "BF16", "FP64_GROUPED"])
def test_emulated_update_within_rounding_bound(arith):
    """Every element of every case, all three functors, with and without
    truncation, ranks 1..128 (both lane layouts)."""
    bad = []
    for n, c in enumerate(_cases(arith, seed=1)):
        w, h = _run(c, arith)
        v = _violations(c, arith, w, h)
        if v:
            bad.append((n, c["kind"], len(c["w"]), c["truncate"], v))
    assert not bad, bad


@pytest.mark.parametrize("arith", ["FP32", "BF16", "FP64_GROUPED"])
@pytest.mark.parametrize("mutant", [1, 2, 3])
def test_rounding_bound_catches_planted_mistakes(arith, mutant):
    """The same bound fails on each planted mistake: 1 drops the
    regularization term, 2 updates h with the new w, 3 swaps nnz1 and nnz2."""
    caught = 0
    for c in _cases(arith, seed=1):
        if mutant in (1, 3) and c["kind"] == "Nzsl":
            continue   # no regularization term to drop or scale
        if mutant == 3 and c["kind"] == "Nzsl_Nzl2":
            continue   # counts unused
        w, h = _run(c, arith, mutant)
        caught += _violations(c, arith, w, h) > 0
    assert caught >= 10, caught


def _oracle_pair(update, truncate, arith, mutant=0, seed=12):
    p = random_problem(400, 500, 20000, 16, seed=seed, values="int")
    res = []
    for a in (None, arith):
        o = po.Oracle(update=update, loss=update, truncate=truncate, rank=p["rank"], tasks_per_rank=4, seed=41,
                      threads=4)
        o.load_coo(0, p["m"], p["n"], *p["train"])
        o.set_factors(p["W0"], p["H0"])
        if a:
            o.set_arith(a, mutant)
        res.append(o)
    return res


EPOCH_RTOL = {"FP32": 1e-6, "BF16": 2e-2, "FP64_GROUPED": 1e-12}
UPDATES = [("Nzsl_L2(0.05)", None), ("Nzsl", None), ("Nzsl_Nzl2(0.02)", None), ("Nzsl_L2(0.05)", "(-0.4,0.4)")]


@pytest.mark.parametrize("arith", ["FP32", "BF16", "FP64_GROUPED"])
@pytest.mark.parametrize("update,truncate", UPDATES)
def test_emulated_epoch_loss_tracks_reference(arith, update, truncate):
    """Five epochs on the same schedules and visit orders: per-epoch training
    loss of the emulation within EPOCH_RTOL of the reference arithmetic's
    (observed: FP32 <= 6e-8, BF16 <= 3.4e-3, FP64 ~1e-16)."""
    ref, emu = _oracle_pair(update, truncate, arith)
    for e in range(5):
        ref.epoch(0.01)
        emu.epoch(0.01)
        lr, le = ref.loss(0), emu.loss(0)
        assert abs(le - lr) <= EPOCH_RTOL[arith] * abs(lr), (e, le, lr, abs(le - lr) / abs(lr))


@pytest.mark.parametrize("arith", ["FP32", "FP64_GROUPED"])
@pytest.mark.parametrize("mutant", [1, 2, 3])
def test_epoch_tolerance_catches_planted_mistakes(arith, mutant):
    """The per-epoch tolerances of FP32 and FP64 fail on each planted mistake
    within five epochs of Nzsl_L2(0.05)."""
    ref, emu = _oracle_pair("Nzsl_L2(0.05)", None, arith, mutant)
    worst = 0.0
    for e in range(5):
        ref.epoch(0.01)
        emu.epoch(0.01)
        lr, le = ref.loss(0), emu.loss(0)
        worst = max(worst, abs(le - lr) / abs(lr))
    assert worst > 10 * EPOCH_RTOL[arith], worst
