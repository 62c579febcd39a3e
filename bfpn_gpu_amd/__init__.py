"""MI355X-native in-core eq15 BFPn solver.

Python view of the C ABI in include/bfpn.h (libbfpn_hip.so, HIP kernels for
gfx950).  There is no CPU fallback: importing the package works anywhere,
but every call that needs the library raises if libbfpn_hip.so is missing or
no HIP device is present.

Main entry points
  Solver(case, precision)   one beam resident in HBM (BFPnSolver)
  ops.*                     operator-level calls on device arrays
  metrics.analyze           BP / P90 / D90 / D20 of an IDD curve
"""
from __future__ import annotations

from .abi import Case, Deltas, Grid, Phys  # noqa: F401
from ._lib import (  # noqa: F401
    FP32,
    FP64,
    BfpnError,
    DeviceArray,
    Server,
    Solver,
    Workspace,
    check,
    device_count,
    lib,
    lib_path,
    stopping_power,
)
from . import metrics, ops, tables  # noqa: F401
