"""sgnx -- MI355X (gfx950) FR/SGN right-hand side and adaptive BS3(2) step.

Python mirror of the reference's C++ solver API over the libsgnx C-ABI
(include/sgnx.h). The numerical work runs in the HIP kernels of libsgnx.so.
"""
from ._lib import (BC_INFLOW, BC_PERIODIC, BC_WALL, LIB_PATH, Params, RkOpts, RkStats, SgnxError,
                   load)
from .solver import H, HP, HU, HV, HW, NVAR, SGN2D, FROperators1D, face_coords, node_coords
from .batch import Batch

AdaptiveRKOptions = RkOpts

__all__ = [
    "SGN2D", "Batch", "FROperators1D", "Params", "RkOpts", "RkStats", "AdaptiveRKOptions", "SgnxError",
    "load", "node_coords", "face_coords", "NVAR", "H", "HU", "HV", "HW", "HP", "BC_PERIODIC",
    "BC_WALL", "BC_INFLOW", "LIB_PATH",
]
