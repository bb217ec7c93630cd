"""sharpclaw_amd: SharpClaw's semi-discrete step (WENO5 + Riemann solve +
SSP Runge-Kutta stage) on MI355X, behind the C ABI of
include/sharpclaw_amd.h (libsharpclaw_amd.so, built by sharpclaw_amd.build).
"""
from ._abi import (PHYSICS, SC_FP32, SC_FP64, SC_PHYS_ACOUSTICS, SC_PHYS_ADVECTION,  # noqa: F401
                   SC_PHYS_EULER, SC_PHYS_SHALLOW_WATER, ScBatchResult, ScParams, ScRequestStats,
                   ScRunCtl, ScRunStats, ScServerStats,
                   grid_shape, interior, make_params)
from ._lib import LIB_PATH, ScError, lib  # noqa: F401
from .serve import Server  # noqa: F401
from .solver import Batch, Solver, copy_bench, evec, flux1, rp, weno5, weno5_char  # noqa: F401
