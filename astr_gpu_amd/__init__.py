"""MI355X (gfx950) FP64 solver for ASTR's per-timestep right-hand side.

The compute path is the HIP library astr_gpu_amd/libastr_hip.so behind the
C ABI of include/astr_hip.h; this package is its host-side binding.
"""
from ._lib import AstrError, AstrParams, load  # noqa: F401
from .solver import (UDF_RIEMANN2D, AstrBatch, AstrSolver, box_metrics, cfl_dt,  # noqa: F401
                     default_params, ensemble_dims, ensemble_member_seed,
                     haloed_shape, hit_state, interior, riemann2d_state,
                     state_from_velocity, tgv_state)

__all__ = ["UDF_RIEMANN2D", "AstrBatch", "AstrError", "AstrParams", "AstrSolver", "box_metrics",
           "cfl_dt", "default_params", "ensemble_dims",
           "ensemble_member_seed", "haloed_shape", "hit_state", "interior", "load", "riemann2d_state",
           "state_from_velocity", "tgv_state"]
