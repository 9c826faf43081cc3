"""dsgdpp_amd -- MI355X-native (gfx950) DSGD++ stratified-SGD epoch.

Drop-in for the mfdsgdpp path of uma-pi1/DSGDpp: C ABI in include/dsgdpp.h,
reference-compatible CLIs in dsgdpp_amd/bin/mfdsgdpp and bin/mfdsgd (DSGD),
Python mirror of the runner/functor interface in dsgdpp_amd.api.
"""
from ._lib import (BALANCE, PRECISIONS, SGD_ORDERS, STRATUM_ORDERS, TEST, TRAIN, LIB_PATH,  # noqa: F401
                   load)
from .api import (BoldDriver, DecayConstant, DsgdJob, DsgdPp, DsgdPpError, DsgdPpJob, DsgdPpRunner,  # noqa: F401
                  DsgdRunner, L2Loss, Nzl2Loss, NzslLoss, RegularizeNone, SumLoss, Trace, TraceEntry,
                  UpdateNzsl, UpdateNzslL2, UpdateNzslNzl2, UpdateTruncate)
