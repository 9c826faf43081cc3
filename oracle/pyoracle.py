"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so (the CPU restatement of the DSGD++ path,
oracle/dsgdpp_oracle.cpp).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use it, as the checker / the timed CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
CLI = os.path.join(HERE, "mfdsgdpp_cpu")
SGD = {"SEQ": 0, "WR": 1, "WOR": 2}
STRATUM = {"SEQ": 0, "RSEQ": 1, "WR": 2, "WOR": 3, "COWOR": 4}
ARITH = {"FP64": 0, "FP32": 1, "BF16": 2, "FP64_GROUPED": 3}
ALGORITHMS = {"DSGDPP": 0, "DSGD": 1}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        P, i32, i64, u64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
        cp = ctypes.c_char_p
        pd = ctypes.POINTER(ctypes.c_double)
        pi32 = ctypes.POINTER(ctypes.c_int32)
        pu32 = ctypes.POINTER(ctypes.c_uint32)
        pi64 = ctypes.POINTER(ctypes.c_int64)
        pu64 = ctypes.POINTER(ctypes.c_uint64)
        sig = {
            "orc_create": (P, [cp, cp, cp, cp, i64, i32, i32, i32, i32, u64, i32, cp, i32]),
            "orc_destroy": (None, [P]),
            "orc_last_error": (cp, [P]),
            "orc_load_coo": (i32, [P, i32, i64, i64, i64, pi32, pi32, pd]),
            "orc_load_mmc": (i32, [P, i32, cp]),
            "orc_load_factors_mma": (i32, [P, cp, cp]),
            "orc_set_factors": (i32, [P, pd, pd]),
            "orc_get_factors": (i32, [P, pd, pd]),
            "orc_write_factors_mma": (i32, [P, cp, cp]),
            "orc_epoch": (i32, [P, dbl, pd]),
            "orc_loss": (i32, [P, i32, pd]),
            "orc_balance": (i32, [P, i32]),
            "orc_get_schedule": (i32, [P, pu32]),
            "orc_set_record": (i32, [P, i32]),
            "orc_set_arith": (i32, [P, i32]),
            "orc_set_algorithm": (i32, [P, i32]),
            "orc_get_permutation": (i32, [P, i32, i32, pu32, i64, pi64]),
            "orc_get_task_seeds": (i32, [P, pu64]),
            "orc_block_nnz": (i32, [P, i32, pi64]),
            "orc_get_block": (i32, [P, i32, i32, i32, pi32, pi32, pd, i64, pi64]),
            "orc_dims": (i32, [P, pi64]),
            "orc_run": (i32, [P, i64, cp, i32, cp, cp, cp, pd, pd, pd, pd, i32]),
            "orc_philox4x32_10": (None, [pu32, pu32, pu32]),
            "orc_rng_draws": (None, [u64, i32, u64, i64, pu64]),
            "orc_dsgd_schedule": (i32, [i32, i32, i32, u64, pu32]),
            "orc_dsgdpp_schedule": (i32, [i32, i32, i32, u64, pu32, pu64]),
            "orc_block_offsets": (None, [i64, i64, pi64]),
            "orc_update_point": (None, [i32, dbl, i64, pd, pd, dbl, dbl, dbl, dbl, dbl, dbl, i32]),
            "orc_update_point_arith": (i32, [i32, i32, i32, dbl, i64, pd, pd, dbl, dbl, dbl, dbl, dbl, dbl, i32]),
            "orc_permute": (None, [u64, u64, i64, pu32, i32]),
            "orc_format_trace": (i32, [i32, pd, pd, i32, pd, pd, pd, pd, pd, cp, cp, ctypes.c_char_p, i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Oracle:
    def __init__(self, update="Nzsl_L2(1)", regularize="None", loss="Nzsl_L2(1)", truncate=None,
                 rank=10, world_size=1, tasks_per_rank=1, sgd_order="WOR", stratum_order="COWOR",
                 seed=0, threads=1, algorithm="DSGDPP"):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self._h = L.orc_create(update.encode(), regularize.encode(), loss.encode(),
                               truncate.encode() if truncate else None, int(rank), int(world_size),
                               int(tasks_per_rank), SGD[sgd_order] if isinstance(sgd_order, str) else sgd_order,
                               STRATUM[stratum_order] if isinstance(stratum_order, str) else stratum_order,
                               int(seed) & ((1 << 64) - 1), int(threads), err, 512)
        if not self._h:
            raise ValueError(err.value.decode())
        self.algorithm = algorithm
        if algorithm != "DSGDPP":
            self._check(L.orc_set_algorithm(self._h, ALGORITHMS[algorithm]))
        self.rank = int(rank)
        self.d = int(world_size) * int(tasks_per_rank)
        self.m = self.n = None

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error(self._h).decode())

    def load_coo(self, which, m, n, i, j, x):
        i = np.ascontiguousarray(i, dtype=np.int32)
        j = np.ascontiguousarray(j, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._check(lib().orc_load_coo(self._h, which, m, n, x.size, _p(i, ctypes.c_int32),
                                       _p(j, ctypes.c_int32), _p(x, ctypes.c_double)))
        self.m, self.n = int(m), int(n)

    def set_factors(self, W, H):
        W = np.ascontiguousarray(W, dtype=np.float64)
        Hcm = np.ascontiguousarray(np.asarray(H, dtype=np.float64).T)
        self._check(lib().orc_set_factors(self._h, _p(W, ctypes.c_double), _p(Hcm, ctypes.c_double)))

    def get_factors(self):
        W = np.empty((self.m, self.rank))
        Hcm = np.empty((self.n, self.rank))
        self._check(lib().orc_get_factors(self._h, _p(W, ctypes.c_double), _p(Hcm, ctypes.c_double)))
        return W, np.ascontiguousarray(Hcm.T)

    def epoch(self, eps):
        s = ctypes.c_double()
        self._check(lib().orc_epoch(self._h, float(eps), ctypes.byref(s)))
        return s.value

    def loss(self, which=0):
        out = ctypes.c_double()
        self._check(lib().orc_loss(self._h, which, ctypes.byref(out)))
        return out.value

    def balance(self, kind):
        self._check(lib().orc_balance(self._h, {"None": 0, "L2": 1, "NZL2": 2}[kind]))

    def set_arith(self, mode, mutant=0):
        """0 FP64 reference; 'FP32'/'BF16'/'FP64_GROUPED': emulate the GPU grouped
        update.  mutant k > 0 plants mistake k in the emulation (see
        update_point_arith); only the tests of the emulation's pins use it."""
        m = ARITH.get(mode, mode)
        self._check(lib().orc_set_arith(self._h, int(m) | (int(mutant) << 4)))

    def record(self, on=True):
        self._check(lib().orc_set_record(self._h, 1 if on else 0))

    @property
    def subepochs(self):
        """Sub-epochs per epoch = column blocks: 2d (DSGD++) or d (DSGD)."""
        return self.d if self.algorithm == "DSGD" else 2 * self.d

    def schedule(self):
        out = np.empty((self.subepochs, self.d), dtype=np.uint32)
        self._check(lib().orc_get_schedule(self._h, _p(out, ctypes.c_uint32)))
        return out

    def task_seeds(self):
        out = np.empty(self.d, dtype=np.uint64)
        self._check(lib().orc_get_task_seeds(self._h, _p(out, ctypes.c_uint64)))
        return out

    def permutation(self, task, subepoch):
        n = ctypes.c_int64()
        self._check(lib().orc_get_permutation(self._h, task, subepoch, None, 0, ctypes.byref(n)))
        out = np.empty(max(1, n.value), dtype=np.uint32)
        self._check(lib().orc_get_permutation(self._h, task, subepoch, _p(out, ctypes.c_uint32),
                                              out.size, ctypes.byref(n)))
        return out[: n.value]

    def block_nnz(self, which=0):
        out = np.empty((self.d, self.subepochs), dtype=np.int64)
        self._check(lib().orc_block_nnz(self._h, which, _p(out, ctypes.c_int64)))
        return out

    def block(self, which, b1, b2):
        n = ctypes.c_int64()
        cap = int(self.block_nnz(which)[b1, b2])
        i = np.empty(max(cap, 1), dtype=np.int32)
        j = np.empty(max(cap, 1), dtype=np.int32)
        x = np.empty(max(cap, 1))
        self._check(lib().orc_get_block(self._h, which, b1, b2, _p(i, ctypes.c_int32), _p(j, ctypes.c_int32),
                                        _p(x, ctypes.c_double), cap, ctypes.byref(n)))
        return i[:cap], j[:cap], x[:cap]

    def run(self, epochs, decay, balance=0, trace_path=None, trace_var="trace", input_file="", verbose=0):
        k = int(epochs) + 1
        arrs = [np.full(k, np.nan) for _ in range(4)]
        self._check(lib().orc_run(self._h, int(epochs), decay.encode(), int(balance),
                                  trace_path.encode() if trace_path else None, trace_var.encode(),
                                  input_file.encode(), *[_p(a, ctypes.c_double) for a in arrs], int(verbose)))
        return dict(zip(("loss", "test_loss", "eps", "time_epoch"), arrs))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def rng_draws(seed, kind, bound, count):
    out = np.empty(count, dtype=np.uint64)
    lib().orc_rng_draws(int(seed), int(kind), int(bound), int(count), _p(out, ctypes.c_uint64))
    return out


def dsgd_schedule(w, t, order, seed):
    d = w * t
    out = np.empty((d, d), dtype=np.uint32)
    rc = lib().orc_dsgd_schedule(w, t, STRATUM[order] if isinstance(order, str) else order, int(seed),
                                 _p(out, ctypes.c_uint32))
    assert rc == 0
    return out


def dsgdpp_schedule(w, t, order, seed):
    d = w * t
    out = np.empty((2 * d, d), dtype=np.uint32)
    seeds = np.empty(d, dtype=np.uint64)
    rc = lib().orc_dsgdpp_schedule(w, t, STRATUM[order] if isinstance(order, str) else order, int(seed),
                                   _p(out, ctypes.c_uint32), _p(seeds, ctypes.c_uint64))
    assert rc == 0
    return out, seeds


def block_offsets(size, blocks):
    out = np.empty(blocks + 1, dtype=np.int64)
    lib().orc_block_offsets(size, blocks, _p(out, ctypes.c_int64))
    return out


def update_point(kind, lam, w, h, x, eps, n1, n2, tmin=0.0, tmax=0.0, truncate=False):
    w = np.array(w, dtype=np.float64)
    h = np.array(h, dtype=np.float64)
    k = {"Nzsl": 0, "Nzsl_L2": 1, "Nzsl_Nzl2": 2}[kind] if isinstance(kind, str) else kind
    lib().orc_update_point(k, float(lam), w.size, _p(w, ctypes.c_double), _p(h, ctypes.c_double), float(x),
                           float(eps), float(n1), float(n2), float(tmin), float(tmax), 1 if truncate else 0)
    return w, h


def update_point_arith(arith, kind, lam, w, h, x, eps, n1, n2, tmin=0.0, tmax=0.0, truncate=False, mutant=0):
    """One update step in arithmetic `arith` ('FP64' = the reference association,
    'FP32' / 'BF16' / 'FP64_GROUPED' = the GPU grouped kernel's); mutant k > 0
    plants mistake k (1 drops the regularization term, 2 updates h with the new
    w, 3 swaps nnz1 and nnz2) for the tests that show the pins catch it."""
    w = np.array(w, dtype=np.float64)
    h = np.array(h, dtype=np.float64)
    k = {"Nzsl": 0, "Nzsl_L2": 1, "Nzsl_Nzl2": 2}[kind] if isinstance(kind, str) else kind
    rc = lib().orc_update_point_arith(ARITH.get(arith, arith), int(mutant), k, float(lam), w.size,
                                      _p(w, ctypes.c_double), _p(h, ctypes.c_double), float(x), float(eps),
                                      float(n1), float(n2), float(tmin), float(tmax), 1 if truncate else 0)
    if rc != 0:
        raise ValueError("invalid arguments to update_point_arith")
    return w, h


def permute(task_seed, first_draw, n, start=None):
    if start is None:
        p = np.arange(n, dtype=np.uint32)
        carry = 0
    else:
        p = np.array(start, dtype=np.uint32)
        carry = 1
    lib().orc_permute(int(task_seed), int(first_draw), int(n), _p(p, ctypes.c_uint32), carry)
    return p


def format_trace(loss, test, eps, t_eps, t_epoch, t_loss, t_test, var="trace", info=""):
    n = len(loss)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (loss, test if test is not None else loss,
                                                               eps, t_eps, t_epoch, t_loss,
                                                               t_test if t_test is not None else t_loss)]
    cap = 1 << 20
    buf = ctypes.create_string_buffer(cap)
    rc = lib().orc_format_trace(n, _p(arrs[0], ctypes.c_double), _p(arrs[1], ctypes.c_double),
                                1 if test is not None else 0, _p(arrs[2], ctypes.c_double),
                                _p(arrs[3], ctypes.c_double), _p(arrs[4], ctypes.c_double),
                                _p(arrs[5], ctypes.c_double), _p(arrs[6], ctypes.c_double), var.encode(),
                                info.encode(), buf, cap)
    assert rc == 0
    return buf.value.decode()
