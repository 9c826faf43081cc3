"""Python mirror of the reference's DSGD++ programming interface.

Reference interface (uma-pi1/DSGDpp):
  * functors        UpdateNzsl / UpdateNzslL2(l) / UpdateNzslNzl2(l) /
                    UpdateTruncate(u, min, max)   (mf/sgd/functions/*.h)
                    RegularizeNone                 (regularize-none.h:22-31)
                    NzslLoss, L2Loss(l), Nzl2Loss(l), SumLoss(a, b)   (mf/loss/*.h)
                    BoldDriver(eps, dec=0.5, inc=1.05), DecayConstant(eps)
                                                   (mf/sgd/decay/*.h)
  * DsgdPpJob(...)  (mf/sgd/dsgdpp.h:39)
  * DsgdPpRunner(seed).run(job, loss, epochs, decay, trace, balance, test)
                    and .epoch(job, eps)           (mf/sgd/dsgdpp.h:82-134)
  * DsgdJob / DsgdRunner, the same for DSGD        (mf/sgd/dsgd.h: d x d blocks)
  * Trace           (mf/trace.h:179-262)

Every device operation goes through the C ABI (include/dsgdpp.h) of the
in-tree HIP library; the step-size decay and the trace stay on the host, as in
the reference's defaultRunner (mf/sgd/sgd_impl.h:34-145).
"""
import ctypes
import math
import sys
import time

import numpy as np

from . import _lib as L


class DsgdPpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# functors (rendered to the reference's functor strings, tools/parse.h:39-61)
# ---------------------------------------------------------------------------
def _num(x):
    """Shortest round-trip decimal ("1", "0.05"), as a user writes it."""
    r = repr(float(x))
    return r[:-2] if r.endswith(".0") else r


class UpdateNzsl:
    def spec(self):
        return "Nzsl"


class UpdateNzslL2:
    def __init__(self, lam):
        self.lam = lam

    def spec(self):
        return f"Nzsl_L2({_num(self.lam)})"


class UpdateNzslNzl2:
    def __init__(self, lam):
        self.lam = lam

    def spec(self):
        return f"Nzsl_Nzl2({_num(self.lam)})"


class UpdateTruncate:
    """UpdateTruncate<U> (update-truncate.h:32-90): clamp to (min+eps, max-eps)."""

    def __init__(self, update, lo, hi):
        self.update, self.lo, self.hi = update, lo, hi

    def spec(self):
        return self.update.spec()

    def truncate(self):
        return f"({_num(self.lo)},{_num(self.hi)})"


class RegularizeNone:
    def spec(self):
        return "None"


class NzslLoss:
    def spec(self):
        return "Nzsl"


class L2Loss:
    def __init__(self, lam):
        self.lam = lam


class Nzl2Loss:
    def __init__(self, lam):
        self.lam = lam


class SumLoss:
    def __init__(self, a, b):
        self.a, self.b = a, b

    def spec(self):
        if isinstance(self.b, L2Loss):
            return f"Nzsl_L2({_num(self.b.lam)})"
        if isinstance(self.b, Nzl2Loss):
            return f"Nzsl_Nzl2({_num(self.b.lam)})"
        raise ValueError("SumLoss supports NzslLoss + L2Loss / Nzl2Loss")


class BoldDriver:
    """decay_bolddriver.h:33-92: eps0 first; then x dec if the loss did not
    decrease (prev <= cur), else x inc.  No rollback of the factors."""

    def __init__(self, eps, decrease=0.5, increase=1.05):
        self.eps, self.decrease, self.increase = float(eps), float(decrease), float(increase)

    def __call__(self, prev, cur):
        if prev is None:
            return self.eps
        self.eps *= self.decrease if prev <= cur else self.increase
        return self.eps

    def spec(self):
        if self.decrease == 0.5 and self.increase == 1.05:
            return f"BoldDriver({_num(self.eps)})"
        return f"BoldDriver({_num(self.eps)},{_num(self.decrease)},{_num(self.increase)})"


class DecayConstant:
    def __init__(self, eps):
        self.eps = float(eps)

    def __call__(self, prev, cur):
        return self.eps

    def spec(self):
        return f"Const({_num(self.eps)})"


# ---------------------------------------------------------------------------
# device context
# ---------------------------------------------------------------------------
class DsgdPp:
    """One DSGD++ problem on one GPU (a dsgdpp_ctx); algorithm="DSGD" runs
    DSGD instead (DsgdRunner, mf/sgd/dsgd_impl.h: d x d blocks, d sub-epochs)."""

    def __init__(self, update="Nzsl_L2(1)", regularize="None", loss="Nzsl_L2(1)", rank=10,
                 tasks_per_rank=1, world_size=1, sgd_order="WOR", stratum_order="COWOR",
                 seed=0, precision="FP64_STRICT", truncate=None, device=0, kernel=0, algorithm="DSGDPP"):
        self._lib = L.load()
        cfg = L.Cfg()
        cfg.abi_version = L.ABI_VERSION
        self._strs = [s.encode() if s is not None else None for s in (update, regularize, loss, truncate)]
        cfg.update, cfg.regularize, cfg.loss, cfg.truncate = self._strs
        cfg.rank = int(rank)
        cfg.world_size = int(world_size)
        cfg.tasks_per_rank = int(tasks_per_rank)
        cfg.sgd_order = L.SGD_ORDERS[sgd_order] if isinstance(sgd_order, str) else int(sgd_order)
        cfg.stratum_order = (L.STRATUM_ORDERS[stratum_order] if isinstance(stratum_order, str)
                             else int(stratum_order))
        cfg.seed = int(seed) & ((1 << 64) - 1)
        cfg.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg.device = int(device)
        cfg.kernel = int(kernel)
        cfg.algorithm = L.ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        h = ctypes.c_void_p()
        rc = self._lib.dsgdpp_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise DsgdPpError(rc, self._lib.dsgdpp_create_error().decode())
        self._h = h
        self.rank = int(rank)
        self.d = int(world_size) * int(tasks_per_rank)
        self.precision = precision
        self.algorithm = algorithm if isinstance(algorithm, str) else {0: "DSGDPP", 1: "DSGD"}[int(algorithm)]
        self.m = self.n = None

    @property
    def subepochs(self):
        """Sub-epochs per epoch = column blocks: 2d (DSGD++) or d (DSGD)."""
        return self.d if self.algorithm == "DSGD" else 2 * self.d

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dsgdpp_destroy(self._h)
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

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise DsgdPpError(rc, self._lib.dsgdpp_last_error(self._h).decode())

    def grouped_ok(self):
        """(True, note) when epochs use the grouped kernel, else (False, reason);
        the note says how many blocks run two-slot groups (empty if none)."""
        ok = ctypes.c_int32()
        buf = ctypes.create_string_buffer(256)
        self._check(self._lib.dsgdpp_grouped_ok(self._h, ctypes.byref(ok), buf, 256))
        return bool(ok.value), buf.value.decode()

    # data ------------------------------------------------------------------
    def load_coo(self, which, m, n, i, j, x):
        i, j, x = L.as_i32(i), L.as_i32(j), L.as_f64(x)
        self._check(self._lib.dsgdpp_load_coo(self._h, which, int(m), int(n), int(x.size),
                                              L.ptr(i, ctypes.c_int32), L.ptr(j, ctypes.c_int32),
                                              L.ptr(x, ctypes.c_double)))
        self.m, self.n = int(m), int(n)

    def load_mmc(self, which, path):
        self._check(self._lib.dsgdpp_load_mmc(self._h, which, str(path).encode()))
        dims = self.dims()
        self.m, self.n = int(dims[0]), int(dims[1])

    def save_blocks(self, which, path):
        """dsgdpp_save_blocks: the loaded matrix in the device layout, to a binary cache file."""
        self._check(self._lib.dsgdpp_save_blocks(self._h, which, str(path).encode()))

    def load_blocks(self, which, path):
        """dsgdpp_load_blocks: a matrix from a binary cache file (in place of load_coo)."""
        self._check(self._lib.dsgdpp_load_blocks(self._h, which, str(path).encode()))
        d = self.dims()
        self.m, self.n = int(d[0]), int(d[1])

    def load_mmc_cached(self, which, path, cache_path):
        """dsgdpp_load_mmc_cached; True when the cache was used."""
        hit = ctypes.c_int32(0)
        self._check(self._lib.dsgdpp_load_mmc_cached(self._h, which, str(path).encode(), str(cache_path).encode(),
                                                     ctypes.byref(hit)))
        d = self.dims()
        self.m, self.n = int(d[0]), int(d[1])
        return bool(hit.value)

    def set_factors(self, W, H):
        """W: m x r row-major; H: r x n (any layout; stored column-major)."""
        W = L.as_f64(W)
        Hcm = np.ascontiguousarray(np.asarray(H, dtype=np.float64).T)   # n x r == column-major r x n
        assert W.shape == (self.m, self.rank) and Hcm.shape == (self.n, self.rank)
        self._check(self._lib.dsgdpp_set_factors(self._h, L.ptr(W, ctypes.c_double),
                                                 L.ptr(Hcm, ctypes.c_double)))

    def get_factors(self):
        W = np.empty((self.m, self.rank), dtype=np.float64)
        Hcm = np.empty((self.n, self.rank), dtype=np.float64)
        self._check(self._lib.dsgdpp_get_factors(self._h, L.ptr(W, ctypes.c_double),
                                                 L.ptr(Hcm, ctypes.c_double)))
        return W, np.ascontiguousarray(Hcm.T)

    def load_factors_mma(self, wpath, hpath):
        self._check(self._lib.dsgdpp_load_factors_mma(self._h, str(wpath).encode(), str(hpath).encode()))

    def write_factors_mma(self, wpath, hpath):
        self._check(self._lib.dsgdpp_write_factors_mma(
            self._h, str(wpath).encode() if wpath else None, str(hpath).encode() if hpath else None))

    # algorithm ---------------------------------------------------------------
    def epoch(self, eps):
        st = L.EpochStats()
        self._check(self._lib.dsgdpp_epoch(self._h, float(eps), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in L.EpochStats._fields_}

    def loss(self, which=L.TRAIN):
        out = ctypes.c_double()
        self._check(self._lib.dsgdpp_loss(self._h, which, ctypes.byref(out)))
        return out.value

    def loss_noalloc(self, which, d_out, stream=0):
        """dsgdpp_loss_noalloc: enqueue the loss on `stream` (a hipStream_t as
        an integer) and write it to the device double at address `d_out`,
        without allocating or waiting."""
        self._check(self._lib.dsgdpp_loss_noalloc(self._h, which, ctypes.c_void_p(int(d_out)),
                                                  ctypes.c_void_p(int(stream))))

    def balance(self, kind="None"):
        self._check(self._lib.dsgdpp_balance(self._h, L.BALANCE[kind] if isinstance(kind, str) else kind))

    def run_native(self, epochs, decay, balance="None", trace_path=None, trace_var="trace",
                   input_file="", verbose=False):
        """The C++ defaultRunner loop (dsgdpp_run)."""
        k = int(epochs) + 1
        arrs = {name: np.full(k, np.nan) for name in ("loss", "test_loss", "eps", "time_epoch")}
        rc = L.RunCfg()
        rc.epochs = int(epochs)
        self._keep = [decay.encode(), trace_path.encode() if trace_path else None, trace_var.encode(),
                      input_file.encode()]
        rc.decay, rc.trace_path, rc.trace_var, rc.input_file = self._keep
        rc.balance = L.BALANCE[balance] if isinstance(balance, str) else int(balance)
        rc.verbose = 1 if verbose else 0
        for name, a in arrs.items():
            setattr(rc, name, L.ptr(a, ctypes.c_double))
        self._check(self._lib.dsgdpp_run(self._h, ctypes.byref(rc)))
        return arrs

    # parity hooks --------------------------------------------------------------
    def schedule(self):
        out = np.empty((self.subepochs, self.d), dtype=np.uint32)
        self._check(self._lib.dsgdpp_get_schedule(self._h, L.ptr(out, ctypes.c_uint32)))
        return out

    def task_seeds(self):
        out = np.empty(self.d, dtype=np.uint64)
        self._check(self._lib.dsgdpp_get_task_seeds(self._h, L.ptr(out, ctypes.c_uint64)))
        return out

    def permutation(self, task, subepoch):
        n = ctypes.c_int64()
        self._check(self._lib.dsgdpp_get_permutation(self._h, task, subepoch, None, 0, ctypes.byref(n)))
        out = np.empty(max(n.value, 1), dtype=np.uint32)
        self._check(self._lib.dsgdpp_get_permutation(self._h, task, subepoch, L.ptr(out, ctypes.c_uint32),
                                                     out.size, ctypes.byref(n)))
        return out[: n.value]

    def block_nnz(self, which=L.TRAIN):
        out = np.empty((self.d, self.subepochs), dtype=np.int64)
        self._check(self._lib.dsgdpp_block_nnz(self._h, which, L.ptr(out, ctypes.c_int64)))
        return out

    def block(self, which, b1, b2):
        """One block's points as stored on the device: (i, j, x) with global indices."""
        n = ctypes.c_int64()
        self._check(self._lib.dsgdpp_get_block(self._h, which, b1, b2, None, None, None, 0, ctypes.byref(n)))
        k = max(n.value, 1)
        i, j, x = np.empty(k, dtype=np.int32), np.empty(k, dtype=np.int32), np.empty(k)
        self._check(self._lib.dsgdpp_get_block(self._h, which, b1, b2, L.ptr(i, ctypes.c_int32),
                                               L.ptr(j, ctypes.c_int32), L.ptr(x, ctypes.c_double), k,
                                               ctypes.byref(n)))
        return i[: n.value], j[: n.value], x[: n.value]

    def dims(self):
        out = np.empty(6, dtype=np.int64)
        self._check(self._lib.dsgdpp_dims(self._h, L.ptr(out, ctypes.c_int64)))
        return out


# ---------------------------------------------------------------------------
# job / runner / trace, as in mf/sgd/dsgdpp.h and mf/trace.h
# ---------------------------------------------------------------------------
class DsgdPpJob:
    """DsgdPpJob (mf/sgd/dsgdpp.h:39): data, factors, functors, orders, d = tasks per rank."""

    algorithm = "DSGDPP"

    def __init__(self, m, n, train, W, H, update, regularize=None, sgd_order="WOR",
                 stratum_order="COWOR", tasks_per_rank=1, loss=None, test=None,
                 precision="FP64_STRICT", seed=0, kernel=0):
        self.update = update
        self.regularize = regularize or RegularizeNone()
        self.loss_spec = loss.spec() if loss is not None else update.spec()
        trunc = update.truncate() if isinstance(update, UpdateTruncate) else None
        self.ctx = DsgdPp(update=update.spec(), regularize=self.regularize.spec(), loss=self.loss_spec,
                          rank=W.shape[1], tasks_per_rank=tasks_per_rank, sgd_order=sgd_order,
                          stratum_order=stratum_order, seed=seed, precision=precision,
                          truncate=trunc, kernel=kernel, algorithm=self.algorithm)
        self.ctx.load_coo(L.TRAIN, m, n, *train)
        if test is not None:
            self.ctx.load_coo(L.TEST, m, n, *test)
        self.ctx.set_factors(W, H)
        self.has_test = test is not None


class DsgdJob(DsgdPpJob):
    """DsgdJob (mf/sgd/dsgd.h): the same job for DSGD -- d x d blocks, one d x d
    schedule per epoch, d sub-epochs (dsgd_impl.h:81-234, 329-350)."""

    algorithm = "DSGD"


class TraceEntry:
    __slots__ = ("it", "loss", "test", "eps", "t_eps", "t_epoch", "t_loss", "t_test")

    def __init__(self, it, loss, test, eps, t_eps, t_epoch, t_loss, t_test):
        self.it, self.loss, self.test, self.eps = it, loss, test, eps
        self.t_eps, self.t_epoch, self.t_loss, self.t_test = t_eps, t_epoch, t_loss, t_test


def _g(v):
    """C++ ostream default formatting of a double (precision 6, %g)."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return "%g" % v


class Trace:
    """mf/trace.h:179-262 (SgdTraceEntry rows; times in nanoseconds)."""

    def __init__(self):
        self.entries = []
        self.dbl = {}
        self.str = {}

    def add(self, e):
        self.entries.append(e)

    def add_field(self, name, value):
        (self.str if isinstance(value, str) else self.dbl)[name] = value

    def to_r(self, var="trace"):
        out = [f"{var} <- list(trace=list("]
        elapsed = 0.0
        for k, e in enumerate(self.entries):
            out.append((", " if k else "") + "\n")
            test = "NA" if (e.test is None or math.isnan(e.test)) else _g(e.test)
            elapsed += (e.t_eps + e.t_epoch + e.t_loss + e.t_test) / 1e9
            eps = "NA" if (e.eps is None or math.isnan(e.eps)) else _g(e.eps)
            out.append(f"list(it={e.it}, epoch={e.it}, loss={_g(e.loss)}, loss.test={test}, "
                       f"time=list(elapsed={_g(elapsed)}, eps={_g(e.t_eps / 1e9)}, "
                       f"epoch={_g(e.t_epoch / 1e9)}, loss={_g(e.t_loss / 1e9)}, "
                       f"loss.test={_g(e.t_test / 1e9)}),  eps={eps})")
        out.append("\n)")
        if self.dbl or self.str:
            items = [f"{k} = {_g(float(v))}" for k, v in sorted(self.dbl.items())]
            items += [f'{k} = "{v}"' for k, v in sorted(self.str.items())]
            out.append(", info=list(" + ", ".join(items) + ")")
        out.append(");")
        return "".join(out)

    def to_rfile(self, path, var="trace"):
        s = self.to_r(var)
        with open(path, "w") as f:
            f.write(s)
        return s


class DsgdPpRunner:
    """DsgdPpRunner (mf/sgd/dsgdpp.h:82-134).  The random source lives in the
    job's context (seeded by the job's seed), as rg::Random32 lives in the
    reference runner."""

    def epoch(self, job, eps):
        return job.ctx.epoch(eps)

    def run(self, job, epochs, decay, trace=None, balance="None", on_epoch=None):
        """defaultRunner (mf/sgd/sgd_impl.h:34-145) with a host-side decay
        object: decay(prev_loss_or_None, cur_loss) -> eps."""
        ctx = job.ctx
        trace = trace if trace is not None else Trace()
        ns = time.perf_counter_ns
        t = ns(); cur = ctx.loss(L.TRAIN); tl = ns() - t
        test, tt = float("nan"), 0
        if job.has_test:
            t = ns(); test = ctx.loss(L.TEST); tt = ns() - t
        trace.add(TraceEntry(0, cur, test, float("nan"), 0, 0, tl, tt))
        prev = None
        for ep in range(int(epochs)):
            t = ns(); eps = decay(prev, cur); te = ns() - t
            t = ns(); stats = ctx.epoch(eps); tep = ns() - t
            if balance != "None":
                ctx.balance(balance)
            prev = cur
            t = ns(); cur = ctx.loss(L.TRAIN); tl = ns() - t
            test, tt = float("nan"), 0
            if job.has_test:
                t = ns(); test = ctx.loss(L.TEST); tt = ns() - t
            trace.add(TraceEntry(ep + 1, cur, test, eps, te, tep, tl, tt))
            if on_epoch is not None:
                on_epoch(ep + 1, cur, test, eps, stats)
        return trace


class DsgdRunner(DsgdPpRunner):
    """DsgdRunner (mf/sgd/dsgd.h): run and epoch as DsgdPpRunner's, on a DsgdJob
    (the same defaultRunner loop, dsgd_impl.h:62-64)."""

    def epoch(self, job, eps):
        if job.algorithm != "DSGD":
            raise TypeError("DsgdRunner runs a DsgdJob")
        return job.ctx.epoch(eps)
