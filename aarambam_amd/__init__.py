"""MI355X-native NG-2LPT initial conditions (drop-in for Aarambam/LPT/Basis).

The compute path is the HIP library libaarambam_hip.so (C-ABI in
include/aarambam.h); this package is its host-side Python binding.
"""
from .api import (AAR_FP32, AAR_FP64, AAR_SCHED_BATCHED, AAR_SCHED_PLAIN, FIELDS, AarError, Context,
                  Job, Server, lib, params_from_dict, read_params, tables)
from .potential import collate_potential

__all__ = ["AAR_FP32", "AAR_FP64", "AAR_SCHED_BATCHED", "AAR_SCHED_PLAIN", "FIELDS", "AarError",
           "Context", "Job", "Server", "collate_potential", "lib", "params_from_dict", "read_params",
           "tables"]
