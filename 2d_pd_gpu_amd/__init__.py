"""MI355X-native explicit bond-associated NOSB peridynamics step (EricYJA/2D_PD_GPU's
computeVelocity loop), exposed through the C ABI of lib/lib2dpdgpu.so.

The package name starts with a digit, so import it with
``importlib.import_module("2d_pd_gpu_amd")``.
"""
from . import pdhip, plates  # noqa: F401
from .pdhip import (BC_FIXED, BC_FREE, FP64, MIXED, Context, Params, PDError,  # noqa: F401
                    build_families)

__all__ = ["pdhip", "plates", "Context", "Params", "PDError", "FP64", "MIXED", "BC_FREE",
           "BC_FIXED", "build_families"]
