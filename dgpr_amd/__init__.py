"""dgpr_amd -- MI355X (gfx950) DG residual and explicit time-step path.

The compute lives in ``libdgx.so`` (hand-written HIP kernels behind the C ABI
declared in ``include/dgx.h``); ``dgpr_amd.dgx`` binds it for Python callers
and ``dgpr_amd.cases`` builds the benchmark workloads.
"""
from . import dgx  # noqa: F401
