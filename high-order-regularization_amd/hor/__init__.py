"""hor - MI355X-native regularized Stokes boundary-integral hot path.

Beale & Tlupova's high-order regularized Stokeslet / stresslet sums, the
barycentric-Lagrange treecode and the two-surface GMRES solve, on one AMD
Instinct MI355X (gfx950) through libhor.so.
"""
from . import geometry, grid, io, ops, solver  # noqa: F401
from ._lib import EXPORTED, HorError, LIB_PATH  # noqa: F401
from .ops import (DIRECT, DL_OFF, DL_ON, FAR_MFMA, FAR_VALU, FP64, MIXED, P_OFF, SL_OFF, SLDL_OFF, SL_ON,  # noqa: F401
                  TREE, Context, Plan, TreeParams, device_count)
