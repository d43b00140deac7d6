"""hydropad_amd -- MI355X-native directionally split PPM hydro step of Hydropad.

The compute path is hand-written gfx950 HIP (csrc/) behind the C-ABI of
include/hydropad_gpu.h; this package is the Python side of that boundary
(ctypes), plus initial-condition generators and the build script.
"""
from . import ic  # noqa: F401
from ._capi import (HP_FP64, HP_MIXED, HpCosmo, HpCounters, HpParams, HydropadError,  # noqa: F401
                    cosmo_init, params_init, timedet_cosmo)
from .hydro import Box, NROT_ORDER, dualenergy_cell, integrate_cell, nrot_of_step, timedet  # noqa: F401

__all__ = ["Box", "HpParams", "HpCounters", "HpCosmo", "HydropadError", "params_init", "timedet",
           "cosmo_init", "timedet_cosmo",
           "nrot_of_step", "NROT_ORDER", "integrate_cell", "dualenergy_cell", "ic",
           "HP_FP64", "HP_MIXED"]
