"""Batched DSGD++: many independent factorizations advance one epoch each in
one persistent launch (C ABI ``dsgdpp_batch_*``), plus the autotuner that
sizes each problem's share of the launch.

Each problem is an ordinary ``api.DsgdPp`` context (its own shape, rank up to
128, lambda, seed and step size); ``DsgdPpBatch.epoch`` runs all of them in
one launch of the grouped kernel, bit-identical to running each context's
``epoch`` alone.  The launch is captured into a HIP graph once and replayed.

Autotuner.  Per problem the free choice is the number of tasks d (d x 2d
blocking, and an H tile of n / 2d columns x 64 or 128 floats in LDS, by
rank).  The launch has at most one workgroup per CU and the library queues
tasks beyond that on the workgroups, so the tuner aims at the device's task
budget: it measures, once per rank class on a probe problem, the update rate
per task for each candidate d (cached on disk), predicts every problem's
epoch time for each candidate, and repeatedly doubles d of the problem with
the largest predicted time while the budget and the problem's shape allow.
Per batch it picks the kernel (``batch_kernel``): the launch's workgroups
take the LDS of the widest problem's tile, and when that leaves no room for
a prep workgroup the batch streams H from HBM instead (kernel 4), so the next
epoch's preps run beside the launch.
"""
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib
from ._lib import TRAIN


class DsgdPpBatch:
    def __init__(self, contexts, use_graph=True):
        self._lib = _lib.load()
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c._h for c in self.contexts])
        h = ctypes.c_void_p()
        rc = self._lib.dsgdpp_batch_create(arr, len(self.contexts), 1 if use_graph else 0, ctypes.byref(h))
        if rc != _lib.OK:
            raise RuntimeError(f"dsgdpp_batch_create failed ({rc}): {self._lib.dsgdpp_batch_create_error().decode()}")
        self._h = h

    def epoch(self, eps):
        eps = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (len(self.contexts),)))
        st = _lib.EpochStats()
        rc = self._lib.dsgdpp_batch_epoch(self._h, _lib.ptr(eps, ctypes.c_double), ctypes.byref(st))
        if rc != _lib.OK:
            raise RuntimeError(f"dsgdpp_batch_epoch failed ({rc}): {self._lib.dsgdpp_batch_last_error(self._h).decode()}")
        return {"seconds": st.seconds, "sgd_seconds": st.sgd_seconds, "updates": st.updates, "launches": st.launches}

    def loss(self, which=TRAIN):
        """Every context's loss (what each context's ``loss(which)`` gives),
        computed on the batch's stream with one wait."""
        out = np.zeros(len(self.contexts), dtype=np.float64)
        rc = self._lib.dsgdpp_batch_loss(self._h, which, _lib.ptr(out, ctypes.c_double))
        if rc != _lib.OK:
            raise RuntimeError(f"dsgdpp_batch_loss failed ({rc}): {self._lib.dsgdpp_batch_last_error(self._h).decode()}")
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dsgdpp_batch_destroy(self._h)
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


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------
CANDIDATE_TASKS = (1, 2, 4, 8, 16, 32)
RANK_CLASSES = (16, 32, 64, 128)
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "dsgdpp_amd", "autotune.json")


def rank_class(r):
    for c in RANK_CLASSES:
        if r <= c:
            return c
    raise ValueError("rank above 128")


def _probe_problem(rank, seed=0):
    """200,000 x 500 with 2M uniform nonzeros, the given rank (narrow enough
    that every candidate d keeps its H tile in LDS)."""
    from . import data
    return data.uniform_problem(200_000, 500, 2_000_000, 0, rank=rank, noise=0.1, seed=seed)


def measure_rates(precision="FP32", update="Nzsl_L2(0.05)", device=0, cache=CACHE, log=None):
    """Updates/s per task for every (rank class, d), measured once per device
    and cached: {rank: {d: rate}}."""
    from . import api
    key = f"{precision}|{update}"
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table:
        return {int(r): {int(d): v for d, v in m.items()} for r, m in table[key].items()}
    rates = {}
    for r in RANK_CLASSES:
        p = _probe_problem(r)
        rates[r] = {}
        for d in CANDIDATE_TASKS:
            g = api.DsgdPp(update=update, loss=update, rank=r, tasks_per_rank=d, seed=1, precision=precision,
                           device=device)
            g.load_coo(TRAIN, p["m"], p["n"], *p["train"])
            if not g.grouped_ok()[0]:   # blocks too large for this d: the planner extrapolates
                g.close()
                continue
            g.set_factors(p["W0"], p["H0"])
            g.epoch(1e-4)   # plans and caches warm
            st = g.epoch(1e-4)
            g.close()
            if st["kernel"] not in (3, 4):
                raise RuntimeError(f"autotune probe rank {r} d {d} did not run the grouped kernel")
            rates[r][d] = len(p["train"][2]) / st["sgd_seconds"] / d
            if log:
                log(f"autotune probe rank {r} d {d}: {rates[r][d] * d:.3g} updates/s")
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        table[key] = {str(r): {str(d): v for d, v in m.items()} for r, m in rates.items()}
        with open(cache, "w") as f:
            json.dump(table, f, indent=1)
    return rates


LDS_BYTES = 160 * 1024
RING_ROWS = 16 * 4   # the kernel's W ring: 16 groups x 4 slots


def row_floats(rank):
    return 64 if rank <= 64 else 128


def lds_need(n, d, rank, wide=False):
    """LDS of one task (bytes).  Tiled: W ring, two H tiles of ceil(n / 2d)
    columns plus the spare column and the columns' factors, two flags.  Wide
    (FP32 column blocks too wide for the tile, H rows streamed from HBM): the
    W and H rings only."""
    if wide:
        return 2 * RING_ROWS * row_floats(rank) * 4
    cols = -(-n // (2 * d))
    tile = ((cols + 1) * row_floats(rank) + cols + 1 + 3) // 4 * 4
    return (RING_ROWS * row_floats(rank) + 2 * tile) * 4 + 8


def prep_lds_estimate(m, n, nnz, d):
    """LDS of the prep kernel for blocks of about 1.5 nnz / 2d^2 points (the
    kernel keeps the visit order and row/column tags, 16-bit below 8,000
    points, plus 4 bits per group, fewer than two groups per point)."""
    cap = block_cap(nnz, d)
    es = 2 if cap < PREP_NARROW else 4
    return (cap + -(-m // d) + -(-n // (2 * d))) * es + cap + 64


PREP_NARROW = 8000   # blocks below this many points take the 16-bit prep


def block_cap(nnz, d):
    """Capacity the planner assumes for the largest of the 2d^2 blocks."""
    return int(1.5 * nnz / (2 * d * d)) + 1


def min_tasks(p, wide=True):
    """Smallest power-of-two d whose prep state fits in LDS and whose H tiles
    fit too, unless the kernel may stream H rows (`wide`).  Blocks are
    further kept below PREP_NARROW points while the shape allows (d <= m,
    2d <= n): the next epoch's prep runs beside the epoch launches, and the
    32-bit prep of larger blocks holds a CU's LDS with one wave, at several
    times the cost per point."""
    d = 1
    while ((not wide and lds_need(p["n"], d, p["rank"]) > LDS_BYTES)
           or prep_lds_estimate(p["m"], p["n"], p["nnz"], d) > LDS_BYTES
           or (block_cap(p["nnz"], d) >= PREP_NARROW and 2 * d <= p["m"] and 4 * d <= p["n"])):
        d *= 2
    return d


def plan_tasks(problems, budget, rates, wide=True):
    """Tasks per problem.  problems: list of dicts with m, n, nnz, rank.
    `wide`: column blocks too wide for the LDS tile may run with H rows
    streamed from HBM (FP32 and BF16).
    Greedy: start at the smallest d the LDS allows (min_tasks); double d of
    the problem with the largest predicted epoch time nnz / (d * rate(d)) while
    the total stays within `budget` and the problem has at least d rows and
    2d columns per block."""
    d = [min_tasks(p, wide) for p in problems]

    def pred(k, dk):
        return predict_time(problems[k], dk, rates)

    while True:
        order = sorted(range(len(problems)), key=lambda k: -pred(k, d[k]))
        grown = False
        for k in order:
            nd = 2 * d[k]
            p = problems[k]
            if nd > max(max(CANDIDATE_TASKS), d[k]) or sum(d) + d[k] > budget or p["m"] < nd or p["n"] < 2 * nd:
                continue
            if pred(k, nd) >= pred(k, d[k]):
                continue
            d[k] = nd
            grown = True
            break
        if not grown:
            return d


def batch_kernel(problems, tasks):
    """Kernel for the batch's contexts: 4 (H rows streamed from HBM, at most
    64 KB of LDS per workgroup) when the widest problem's LDS tile leaves a CU
    too little LDS for the largest prep workgroup beside it, else 0 (the
    tile where it fits).  The preps of the next epoch start in order of LDS
    on one stream; with no room beside the launch they all wait for it
    (batch64 on MI355X: 251 ms per epoch on the tile against 217.7 ms
    streaming, whose launch takes 150 against 136 ms)."""
    def own(p, d):   # a problem whose tile does not fit streams anyway
        t = lds_need(p["n"], d, p["rank"])
        return t if t <= LDS_BYTES else lds_need(p["n"], d, p["rank"], wide=True)

    tile = max(own(p, d) for p, d in zip(problems, tasks))
    # a prep whose state exceeds a CU's LDS keeps it in global memory
    preps = [prep_lds_estimate(p["m"], p["n"], p["nnz"], d) for p, d in zip(problems, tasks)]
    prep = max([q for q in preps if q <= LDS_BYTES], default=0)
    return 4 if tile + prep > LDS_BYTES else 0


def predict_time(p, d, rates):
    """Predicted epoch time (s) of problem p on d tasks: nnz / (d * per-task
    rate), the rate measured at the largest candidate d not above d (the
    smallest candidate below the measured range)."""
    r = rates[rank_class(p["rank"])]
    rate = (r.get(d) or r[max(x for x in r if x <= d)]) if d >= min(r) else r[min(r)]
    return p["nnz"] / (d * rate)


def partition(tasks, budget, times=None):
    """Problem indices split into launches of at most `budget` tasks each;
    a problem larger than the budget runs alone.  First fit, in decreasing
    order of `times` when given (a launch lasts as long as its slowest
    problem, so problems of like time share launches), else of tasks."""
    key = times if times is not None else tasks
    order = sorted(range(len(tasks)), key=lambda k: -key[k])
    parts, load = [], []
    for k in order:
        for q in range(len(parts)):
            if load[q] + tasks[k] <= budget:
                parts[q].append(k)
                load[q] += tasks[k]
                break
        else:
            parts.append([k])
            load.append(tasks[k])
    return [sorted(p) for p in parts]


def batch_problems(count=64, seed=0, shrink=1, max_nnz=20_000_000, log=None):
    """Independent problems: rows log-uniform in [1e4, 1e6], columns
    log-uniform in [1e3, 1e5], nnz = U(1e-4, 1e-2) * rows * cols capped at
    max_nnz, rank-10 N(0,1) ground truth plus N(0, 0.1^2) noise, rank drawn
    from {16, 32, 64, 128}, lambda log-uniform in [1e-3, 1e-1].  `shrink`
    divides rows, columns and the cap (quick runs)."""
    from . import data
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        m = int(math.exp(rng.uniform(math.log(1e4), math.log(1e6))) / shrink)
        n = int(math.exp(rng.uniform(math.log(1e3), math.log(1e5))) / shrink)
        nnz = int(min(rng.uniform(1e-4, 1e-2) * m * n, max_nnz / shrink))
        nnz = max(nnz, 4 * (m + n) // 10)
        nnz = min(nnz, m * n // 2)
        rank = int(rng.choice([16, 32, 64, 128]))
        lam = float(math.exp(rng.uniform(math.log(1e-3), math.log(1e-1))))
        p = data.uniform_problem(m, n, nnz, 0, rank=rank, truth_rank=10, noise=0.1, seed=seed * 1000 + k,
                                 init=0.1)
        p["lambda"] = lam
        out.append(p)
        if log and (k + 1) % 8 == 0:
            log(f"[batch] {k + 1}/{count} problems generated")
    return out


def now():
    return time.perf_counter()
