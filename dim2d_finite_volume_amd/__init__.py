"""MI355X (gfx950) implementation of the DIM time-derivative and RK3-TVD hot
path of jdesmarais/dim2D_finite_volume.

The compute runs in the HIP library lib/libdim2d_hip.so behind the C ABI of
include/dim2d_hip.h; this package is its Python host binding (used by the
tests and the benchmark).  The Fortran binding lives in fortran/.
"""
from . import capi
from .capi import (BC_OPEN, BC_PERIODIC, BC_REFLECTION, BC_WALL, INT_LSRK3, INT_RK3TVD,
                   OPS_COCKBURN_GAU, OPS_MATTSSON, PREC_FP64, PREC_MIXED, Dim2dError)
from .solver import Batch, Field, ModelParams

__all__ = ["capi", "Batch", "Field", "ModelParams", "Dim2dError", "BC_PERIODIC",
           "BC_REFLECTION", "BC_WALL", "BC_OPEN", "PREC_FP64", "PREC_MIXED", "INT_RK3TVD", "INT_LSRK3",
           "OPS_MATTSSON", "OPS_COCKBURN_GAU"]
