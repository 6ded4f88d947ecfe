"""FluidMedium (src/fluid.rs:687-788): the renderer-side view of a solver's
density and temperature, as device snapshots behind the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .solver import _dptr


class FluidMedium:
    """from_solver snapshots density and temperature on the device
    (src/fluid.rs:703-735): later solver steps leave the medium unchanged.
    bounds = ((xmin, ymin, zmin), (xmax, ymax, zmax))."""

    def __init__(self, solver, bounds, sigma_t_scale: float, phase_g: float):
        self._lib = L.lib()
        lo = np.ascontiguousarray(np.asarray(bounds[0], dtype=np.float64))
        hi = np.ascontiguousarray(np.asarray(bounds[1], dtype=np.float64))
        h = C.c_void_p()
        L.check(self._lib.aether_fluid_medium_create(solver._h, _dptr(lo), _dptr(hi),
                                                     float(sigma_t_scale), float(phase_g),
                                                     C.byref(h)))
        self._h = h
        self.bounds = (tuple(lo), tuple(hi))
        self.nx, self.ny, self.nz = solver.nx, solver.ny, solver.nz
        self.sigma_t_scale = float(sigma_t_scale)
        md, mj, g = C.c_double(), C.c_double(), C.c_double()
        L.check(self._lib.aether_fluid_medium_info(h, C.byref(md), C.byref(mj), C.byref(g)))
        self.max_density, self.majorant_sigma_t, self.phase_g = md.value, mj.value, g.value

    @classmethod
    def from_solver(cls, solver, bounds, sigma_t_scale: float, phase_g: float) -> "FluidMedium":
        """src/fluid.rs:703"""
        return cls(solver, bounds, sigma_t_scale, phase_g)

    def sample_density_and_temp(self, points) -> np.ndarray:
        """src/fluid.rs:739-788 at every row of `points` (n x 3); returns an
        (n, 2) array of (density, temperature)."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        out = np.empty((len(pts), 2), dtype=np.float64)
        L.check(self._lib.aether_fluid_medium_sample(self._h, _dptr(pts), len(pts), _dptr(out)))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.aether_fluid_medium_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
