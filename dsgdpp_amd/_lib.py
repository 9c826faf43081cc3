"""ctypes binding of the C ABI declared in include/dsgdpp.h.

The shared library is built in-tree (``dsgdpp_amd/libdsgdpp_amd.so``) by
``__graft_entry__.build()``.  There is no fallback: if the library is missing
or fails to load, importing the GPU path raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdsgdpp_amd.so")

ABI_VERSION = 1

OK, E_ARG, E_STATE, E_IO, E_DEVICE, E_UNSUPPORTED = 0, -1, -2, -3, -4, -5
SGD_ORDERS = {"SEQ": 0, "WR": 1, "WOR": 2}
STRATUM_ORDERS = {"SEQ": 0, "RSEQ": 1, "WR": 2, "WOR": 3, "COWOR": 4}
PRECISIONS = {"FP64_STRICT": 0, "FP64": 1, "FP32": 2, "BF16": 3}
BALANCE = {"None": 0, "L2": 1, "NZL2": 2}
ALGORITHMS = {"DSGDPP": 0, "DSGD": 1}
TRAIN, TEST = 0, 1

# every symbol include/dsgdpp.h declares
EXPORTS = (
    "dsgdpp_create", "dsgdpp_destroy", "dsgdpp_last_error", "dsgdpp_create_error",
    "dsgdpp_load_coo", "dsgdpp_load_mmc", "dsgdpp_set_factors", "dsgdpp_get_factors",
    "dsgdpp_load_factors_mma", "dsgdpp_write_factors_mma", "dsgdpp_epoch", "dsgdpp_loss",
    "dsgdpp_loss_noalloc", "dsgdpp_balance", "dsgdpp_run", "dsgdpp_get_schedule", "dsgdpp_get_task_seeds",
    "dsgdpp_get_permutation", "dsgdpp_block_nnz", "dsgdpp_dims", "dsgdpp_grouped_ok", "dsgdpp_get_block",
    "dsgdpp_save_blocks", "dsgdpp_load_blocks", "dsgdpp_load_mmc_cached",
    "dsgdpp_mlp_num_params", "dsgdpp_mlp_create", "dsgdpp_mlp_create_error", "dsgdpp_mlp_last_error",
    "dsgdpp_mlp_destroy", "dsgdpp_mlp_set_params", "dsgdpp_mlp_get_params", "dsgdpp_mlp_forward",
    "dsgdpp_mlp_gradient", "dsgdpp_mlp_train",
    "dsgdpp_batch_create", "dsgdpp_batch_create_error", "dsgdpp_batch_last_error", "dsgdpp_batch_destroy",
    "dsgdpp_batch_epoch", "dsgdpp_batch_loss",
)


class Cfg(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("update", ctypes.c_char_p),
        ("regularize", ctypes.c_char_p),
        ("loss", ctypes.c_char_p),
        ("truncate", ctypes.c_char_p),
        ("rank", ctypes.c_int64),
        ("world_size", ctypes.c_int32),
        ("tasks_per_rank", ctypes.c_int32),
        ("sgd_order", ctypes.c_int32),
        ("stratum_order", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("kernel", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class EpochStats(ctypes.Structure):
    _fields_ = [
        ("seconds", ctypes.c_double),
        ("device_seconds", ctypes.c_double),
        ("perm_seconds", ctypes.c_double),
        ("sgd_seconds", ctypes.c_double),
        ("updates", ctypes.c_int64),
        ("launches", ctypes.c_int64),
        ("kernel", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class RunCfg(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int64),
        ("decay", ctypes.c_char_p),
        ("balance", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("trace_path", ctypes.c_char_p),
        ("trace_var", ctypes.c_char_p),
        ("input_file", ctypes.c_char_p),
        ("loss", ctypes.POINTER(ctypes.c_double)),
        ("test_loss", ctypes.POINTER(ctypes.c_double)),
        ("eps", ctypes.POINTER(ctypes.c_double)),
        ("time_epoch", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def load():
    """Load the in-tree library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(the GPU path has no fallback)")
    # torch wheels carry their own HIP runtime (torch/lib/libamdhip64.so and
    # libhsa-runtime64.so, asked for by unversioned names).  Loaded after this
    # library's runtime it would be a second runtime in the process, and it
    # finds no device; loaded first, this library binds to it (the sonames
    # match).  So torch, when installed, goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    pi32 = ctypes.POINTER(ctypes.c_int32)
    pu32 = ctypes.POINTER(ctypes.c_uint32)
    pu64 = ctypes.POINTER(ctypes.c_uint64)
    pi64 = ctypes.POINTER(ctypes.c_int64)
    pd = ctypes.POINTER(ctypes.c_double)
    pf = ctypes.POINTER(ctypes.c_float)
    sig = {
        "dsgdpp_create": (i32, [ctypes.POINTER(Cfg), ctypes.POINTER(P)]),
        "dsgdpp_destroy": (None, [P]),
        "dsgdpp_last_error": (ctypes.c_char_p, [P]),
        "dsgdpp_create_error": (ctypes.c_char_p, []),
        "dsgdpp_load_coo": (i32, [P, i32, i64, i64, i64, pi32, pi32, pd]),
        "dsgdpp_load_mmc": (i32, [P, i32, ctypes.c_char_p]),
        "dsgdpp_set_factors": (i32, [P, pd, pd]),
        "dsgdpp_get_factors": (i32, [P, pd, pd]),
        "dsgdpp_load_factors_mma": (i32, [P, ctypes.c_char_p, ctypes.c_char_p]),
        "dsgdpp_write_factors_mma": (i32, [P, ctypes.c_char_p, ctypes.c_char_p]),
        "dsgdpp_epoch": (i32, [P, dbl, ctypes.POINTER(EpochStats)]),
        "dsgdpp_loss": (i32, [P, i32, pd]),
        "dsgdpp_loss_noalloc": (i32, [P, i32, P, P]),
        "dsgdpp_balance": (i32, [P, i32]),
        "dsgdpp_run": (i32, [P, ctypes.POINTER(RunCfg)]),
        "dsgdpp_get_schedule": (i32, [P, pu32]),
        "dsgdpp_get_task_seeds": (i32, [P, pu64]),
        "dsgdpp_get_permutation": (i32, [P, i32, i32, pu32, i64, pi64]),
        "dsgdpp_block_nnz": (i32, [P, i32, pi64]),
        "dsgdpp_get_block": (i32, [P, i32, i32, i32, pi32, pi32, pd, i64, pi64]),
        "dsgdpp_save_blocks": (i32, [P, i32, ctypes.c_char_p]),
        "dsgdpp_load_blocks": (i32, [P, i32, ctypes.c_char_p]),
        "dsgdpp_load_mmc_cached": (i32, [P, i32, ctypes.c_char_p, ctypes.c_char_p, pi32]),
        "dsgdpp_dims": (i32, [P, pi64]),
        "dsgdpp_grouped_ok": (i32, [P, pi32, ctypes.c_char_p, i64]),
        "dsgdpp_mlp_num_params": (i32, []),
        "dsgdpp_mlp_create": (i32, [i32, ctypes.POINTER(P)]),
        "dsgdpp_mlp_create_error": (ctypes.c_char_p, []),
        "dsgdpp_mlp_last_error": (ctypes.c_char_p, [P]),
        "dsgdpp_mlp_destroy": (None, [P]),
        "dsgdpp_mlp_set_params": (i32, [P, pf]),
        "dsgdpp_mlp_get_params": (i32, [P, pf]),
        "dsgdpp_mlp_forward": (i32, [P, pf, i64, pf]),
        "dsgdpp_mlp_gradient": (i32, [P, pf, pf, i64, pf, pd]),
        "dsgdpp_mlp_train": (i32, [P, pf, pf, i64, i64, ctypes.c_float, pd]),
        "dsgdpp_batch_create": (i32, [ctypes.POINTER(P), i32, i32, ctypes.POINTER(P)]),
        "dsgdpp_batch_create_error": (ctypes.c_char_p, []),
        "dsgdpp_batch_last_error": (ctypes.c_char_p, [P]),
        "dsgdpp_batch_destroy": (None, [P]),
        "dsgdpp_batch_epoch": (i32, [P, pd, ctypes.POINTER(EpochStats)]),
        "dsgdpp_batch_loss": (i32, [P, i32, pd]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a, ctype):
    """Pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
