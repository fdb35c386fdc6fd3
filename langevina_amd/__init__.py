"""MI355X-native ModelG hot path (langevinA's PV2HBSplit time step).

The compute path is liblgv_hip.so (HIP kernels for gfx950 behind the C ABI
of include/lgv.h); this package is its thin host binding.
"""
from .lgv import (FP32, FP64, PHI_AUTO, PHI_MARCH, PHI_SITE, Context,  # noqa: F401
                  LgvError, ModelParams, lib, model_params)

__all__ = ["Context", "ModelParams", "model_params", "FP32", "FP64", "PHI_AUTO", "PHI_SITE",
           "PHI_MARCH", "LgvError", "lib"]
