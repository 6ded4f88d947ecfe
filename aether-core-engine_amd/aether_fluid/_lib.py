"""ctypes declarations of include/aether_fluid.h and the library loader.

The shared library ``libaether_fluid_hip.so`` is built in-tree (``lib/`` of
this package's parent directory) by ``build()``.  There is no fallback: if the
library is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("AETHER_FLUID_LIB") or os.path.join(PKG_ROOT, "lib",
                                                                "libaether_fluid_hip.so")

# status codes and enums (include/aether_fluid.h)
AETHER_OK = 0
AETHER_EINVAL = -1
AETHER_EHIP = -2
AETHER_ENOMEM = -3
AETHER_ENODEV = -4
AETHER_EINTERNAL = -5
ABI_VERSION = 3  # include/aether_fluid.h AETHER_FLUID_ABI_VERSION

FP32, FP64, FP32_IR64 = 0, 1, 2
SOLVER_CG, SOLVER_RBGS = 0, 1
SOURCE_JETS, SOURCE_SPHERE, SOURCE_NONE = 0, 1, 2
CG_AUTO, CG_SCALAR, CG_X4, CG_FUSED = 0, 1, 2, 3
FIELDS = {"u": 0, "v": 1, "w": 2, "density": 3, "temperature": 4, "pressure": 5,
          "divergence": 6}


class Params(C.Structure):
    _fields_ = [
        ("vorticity_eps", C.c_double),
        ("buoyancy_beta", C.c_double),
        ("gravity_alpha", C.c_double),
        ("solver_kind", C.c_int32),
        ("solver_iters", C.c_int32),
        ("sor_omega", C.c_double),
        ("cg_rel_tol", C.c_double),
        ("dissipation_velocity", C.c_double),
        ("dissipation_density", C.c_double),
        ("dissipation_temperature", C.c_double),
        ("source_kind", C.c_int32),
        ("cg_form", C.c_int32),
        ("source_center", C.c_double * 3),
        ("source_radius", C.c_double),
        ("source_density", C.c_double),
        ("source_temperature", C.c_double),
        ("ir_inner_rel_tol", C.c_double),
        ("ir_div_tol", C.c_double),
        ("ir_max_outer", C.c_int32),
        ("batch_graph", C.c_int32),
        ("cg_z_cells", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Telemetry(C.Structure):
    _fields_ = [
        ("time_sec", C.c_double),
        ("kinetic_energy", C.c_double),
        ("enstrophy", C.c_double),
        ("peak_vorticity", C.c_double),
        ("total_mass", C.c_double),
        ("center_of_mass", C.c_double * 3),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("b_norm2", C.c_double),
        ("r_norm2_initial", C.c_double),
        ("r_norm2_final", C.c_double),
        ("outer_iterations", C.c_int32),
        ("reserved", C.c_int32),
        ("div_max_final", C.c_double),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
        ("h", C.c_double),
        ("vorticity_eps", C.c_double), ("buoyancy_beta", C.c_double),
        ("gravity_alpha", C.c_double),
        ("source_center", C.c_double * 3), ("source_radius", C.c_double),
        ("source_density", C.c_double), ("source_temperature", C.c_double),
    ]


class RolloutDesc(C.Structure):
    _fields_ = [
        ("nsteps", C.c_int32), ("checkpoint_every", C.c_int32), ("dt", C.c_double),
        ("patch_lo", C.c_int32 * 3), ("patch_hi", C.c_int32 * 3),
        ("adjoint_cg_iters", C.c_int32), ("reserved", C.c_int32),
        ("adjoint_cg_rel_tol", C.c_double),
    ]


# aether_medium_ray and aether_medium_interaction as numpy record layouts
RAY_DTYPE = np.dtype([("origin", "<f8", (3,)), ("dir", "<f8", (3,)), ("t_min", "<f8"),
                      ("t_far", "<f8")])
INTERACTION_DTYPE = np.dtype([("hit", "<i4"), ("reserved", "<i4"), ("t", "<f8"),
                              ("point", "<f8", (3,)), ("albedo", "<f8", (4,)),
                              ("emission", "<f8", (4,))])
assert RAY_DTYPE.itemsize == 64 and INTERACTION_DTYPE.itemsize == 104

# every symbol declared in include/aether_fluid.h: name -> (restype, argtypes)
_vp = C.c_void_p
_pd = C.POINTER(C.c_double)
_pu64 = C.POINTER(C.c_uint64)
SIGNATURES = {
    "aether_fluid_abi_version": (C.c_int, []),
    "aether_fluid_last_error": (C.c_char_p, []),
    "aether_fluid_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "aether_fluid_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                      C.POINTER(_vp)]),
    "aether_fluid_destroy": (C.c_int, [_vp]),
    "aether_fluid_dims": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), _pd, C.POINTER(C.c_int)]),
    "aether_fluid_default_params": (C.c_int, [C.POINTER(Params)]),
    "aether_fluid_set_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "aether_fluid_get_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "aether_fluid_step": (C.c_int, [_vp, C.c_double, C.c_double, C.POINTER(Telemetry)]),
    "aether_fluid_run": (C.c_int, [_vp, C.c_double, _pd, C.c_int, C.POINTER(Telemetry)]),
    "aether_fluid_advect_maccormack_in_place": (C.c_int, [_vp, C.c_int, C.c_double,
                                                          C.c_double]),
    "aether_fluid_enforce_wall_velocity_bc": (C.c_int, [_vp]),
    "aether_fluid_apply_buoyancy_and_vorticity_confinement": (C.c_int, [_vp, C.c_double]),
    "aether_fluid_compute_divergence_norm": (C.c_int, [_vp, _pd]),
    "aether_fluid_project": (C.c_int, [_vp]),
    "aether_fluid_solve_poisson_conjugate_gradient": (C.c_int, [_vp, C.c_int]),
    "aether_fluid_solve_poisson_gauss_seidel": (C.c_int, [_vp, C.c_int]),
    "aether_fluid_inject_sources": (C.c_int, [_vp, C.c_double, C.c_double]),
    "aether_fluid_compute_telemetry": (C.c_int, [_vp, C.c_double, C.POINTER(Telemetry)]),
    "aether_fluid_compute_telemetry_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int,
                                                         C.c_double, _vp, _vp, _vp, _vp,
                                                         C.c_double, _vp, _vp]),
    "aether_fluid_sample_trilinear": (C.c_int, [_vp, C.c_int, _pd, C.c_size_t, _pd]),
    "aether_fluid_sample_trilinear_field": (C.c_int, [_vp, _pd, C.c_size_t, _pd, C.c_size_t,
                                                      _pd]),
    "aether_fluid_divergence_max": (C.c_int, [_vp, _pd]),
    "aether_fluid_last_solve_stats": (C.c_int, [_vp, C.POINTER(SolveStats)]),
    "aether_fluid_last_run_iterations": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_int,
                                                   C.POINTER(C.c_int)]),
    "aether_fluid_upload": (C.c_int, [_vp, C.c_int, _pd, C.c_size_t]),
    "aether_fluid_nonfinite": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "aether_fluid_download": (C.c_int, [_vp, C.c_int, _pd, C.c_size_t]),
    "aether_fluid_max_density": (C.c_int, [_vp, _pd]),
    "aether_fluid_batch_autotune_cg": (C.c_int, [_vp, C.c_int, _pd, C.POINTER(C.c_int),
                                                 C.POINTER(C.c_int)]),
    "aether_fluid_autotune_cg": (C.c_int, [_vp, C.c_int, _pd, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]),
    "aether_fluid_medium_create": (C.c_int, [_vp, _pd, _pd, C.c_double, C.c_double,
                                             C.POINTER(_vp)]),
    "aether_fluid_medium_destroy": (C.c_int, [_vp]),
    "aether_fluid_medium_info": (C.c_int, [_vp, _pd, _pd, _pd]),
    "aether_fluid_medium_sample": (C.c_int, [_vp, _pd, C.c_size_t, _pd]),
    "aether_fluid_medium_sample_interaction": (C.c_int, [_vp, _vp, C.c_size_t, _pd, _pu64,
                                                         _vp]),
    "aether_fluid_medium_transmittance": (C.c_int, [_vp, _vp, C.c_size_t, _pu64, _pd]),
    "aether_fluid_medium_phase_hg_pdf": (C.c_int, [_vp, _pd, _pd, C.c_size_t, _pd]),
    "aether_fluid_medium_sample_phase_hg": (C.c_int, [_vp, _pd, C.c_size_t, _pu64, _pd, _pd]),
    "aether_fluid_batch_create": (C.c_int, [C.POINTER(InstanceDesc), C.c_int, C.c_int,
                                            C.POINTER(Params), C.POINTER(_vp)]),
    "aether_fluid_batch_destroy": (C.c_int, [_vp]),
    "aether_fluid_batch_upload": (C.c_int, [_vp, C.c_int, C.c_int, _pd, C.c_size_t]),
    "aether_fluid_batch_download": (C.c_int, [_vp, C.c_int, C.c_int, _pd, C.c_size_t]),
    "aether_fluid_batch_run": (C.c_int, [_vp, C.c_double, _pd, C.c_int, C.POINTER(Telemetry)]),
    "aether_fluid_batch_iterations": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_int,
                                                C.POINTER(C.c_int)]),
    "aether_fluid_rollout": (C.c_int, [_vp, C.POINTER(RolloutDesc), _pd, C.c_double, _pd, _pd,
                                       _pd, _pd]),
    "aether_fluid_rollout_forward": (C.c_int, [_vp, C.POINTER(RolloutDesc), _pd, C.c_double,
                                               _pd]),
    "aether_fluid_rollout_backward": (C.c_int, [_vp, _pd, _pd, _pd, _pd]),
    "aether_rng_seed": (C.c_int, [C.c_uint64, _pu64]),
    "aether_rng_fill_range": (C.c_int, [_pu64, C.c_double, C.c_double, _pd, C.c_size_t]),
    "aether_fluid_set_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "aether_fluid_stream_copy": (C.c_int, [C.c_size_t, C.c_int, _pd]),
    "aether_fluid_kernel_timing": (C.c_int, [_vp, C.c_int, _pd, C.POINTER(C.c_long)]),
}


class AetherFluidError(RuntimeError):
    """A negative status from the C ABI, with the library's error text."""

    def __init__(self, code: int, message: str):
        super().__init__(f"aether_fluid error {code}: {message}")
        self.code = code


_lib = None


def build() -> None:
    """Compile libaether_fluid_hip.so for gfx950 with the package Makefile."""
    subprocess.run(["make", "-s", "-C", PKG_ROOT], check=True)


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AetherFluidError(AETHER_ENODEV,
                               f"{LIB_PATH} is missing: build it with aether_fluid.build()")
    h = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    if h.aether_fluid_abi_version() != ABI_VERSION:  # the structs below follow version 3
        raise AetherFluidError(AETHER_EINTERNAL,
                               f"{LIB_PATH} has ABI version {h.aether_fluid_abi_version()}, "
                               f"this binding expects {ABI_VERSION}: rebuild it")
    _lib = h
    return h


def check(status: int) -> None:
    if status != AETHER_OK:
        msg = lib().aether_fluid_last_error().decode(errors="replace")
        raise AetherFluidError(status, msg)
