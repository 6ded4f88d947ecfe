"""MI355X (gfx950) smoke solver: a drop-in for the reference's `pub mod fluid`
(src/lib.rs:4, src/fluid.rs).  The compute path is the HIP library
``libaether_fluid_hip.so`` behind include/aether_fluid.h; this package is its
host-side mirror."""
from ._lib import (AetherFluidError, CG_AUTO, CG_FUSED, CG_SCALAR, CG_X4, FIELDS, FP32,
                   FP32_IR64, FP64, SOLVER_CG, SOLVER_RBGS, SOURCE_JETS, SOURCE_NONE,
                   SOURCE_SPHERE, LIB_PATH, SIGNATURES, build, lib)
from .solver import (FluidSolver3D, FluidTelemetryFrame, PoissonSolverKind, rng_fill,
                     stream_copy_ms)
from .scenarios import seeded_config_state, SMOKE_CONFIGS, batched_instances
from .batch import FluidBatch
from .adjoint import SmokeRollout
from .medium import FluidMedium

__all__ = [
    "AetherFluidError", "CG_AUTO", "CG_FUSED", "CG_SCALAR", "CG_X4", "FIELDS", "FP32",
    "FP32_IR64", "FP64", "SOLVER_CG", "SOLVER_RBGS", "SOURCE_JETS", "SOURCE_NONE",
    "SOURCE_SPHERE", "LIB_PATH", "SIGNATURES", "build", "lib", "FluidSolver3D",
    "FluidTelemetryFrame", "PoissonSolverKind", "rng_fill", "stream_copy_ms",
    "seeded_config_state", "SMOKE_CONFIGS", "batched_instances", "FluidBatch", "SmokeRollout",
    "FluidMedium",
]
