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

    # --- volumetric sampling for the path tracer (src/fluid.rs:790-910) ----
    @staticmethod
    def make_rays(origins, dirs, t_min, t_far) -> np.ndarray:
        """Rays as Ray::new gives them: dirs must already be normalised
        (Ray::dir); t_far is the call's surface_t or max_dist."""
        o = np.asarray(origins, dtype=np.float64).reshape(-1, 3)
        r = np.zeros(len(o), dtype=L.RAY_DTYPE)
        r["origin"], r["dir"] = o, np.asarray(dirs, dtype=np.float64).reshape(-1, 3)
        r["t_min"], r["t_far"] = t_min, t_far
        return r

    @staticmethod
    def _states(states) -> np.ndarray:
        s = np.asarray(states)
        if s.dtype != np.uint64 or not s.flags.c_contiguous or s.ndim != 2 or s.shape[1] != 4:
            raise ValueError("rng states must be a C-contiguous (n, 4) uint64 array")
        return s

    def sample_interaction(self, rays, lambdas, states) -> np.ndarray:
        """src/fluid.rs:793-838 for every ray (a RAY_DTYPE array); lambdas
        (n, 4) nm; states (n, 4) uint64 Xoshiro256++ states, advanced in
        place.  Returns an INTERACTION_DTYPE array (hit = 0: None)."""
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64).reshape(-1, 4))
        st = self._states(states)
        assert len(lam) == len(rays) == len(st)
        out = np.zeros(len(rays), dtype=L.INTERACTION_DTYPE)
        L.check(self._lib.aether_fluid_medium_sample_interaction(
            self._h, rays.ctypes.data, len(rays), _dptr(lam),
            st.ctypes.data_as(C.POINTER(C.c_uint64)), out.ctypes.data))
        return out

    def transmittance(self, rays, states) -> np.ndarray:
        """src/fluid.rs:841-874 for every ray; states advanced in place."""
        rays = np.ascontiguousarray(rays, dtype=L.RAY_DTYPE)
        st = self._states(states)
        assert len(rays) == len(st)
        out = np.empty(len(rays), dtype=np.float64)
        L.check(self._lib.aether_fluid_medium_transmittance(
            self._h, rays.ctypes.data, len(rays), st.ctypes.data_as(C.POINTER(C.c_uint64)),
            _dptr(out)))
        return out

    def phase_hg_pdf(self, wo, wi) -> np.ndarray:
        """src/fluid.rs:878-888 for every row pair."""
        wo = np.ascontiguousarray(np.asarray(wo, dtype=np.float64).reshape(-1, 3))
        wi = np.ascontiguousarray(np.asarray(wi, dtype=np.float64).reshape(-1, 3))
        out = np.empty(len(wo), dtype=np.float64)
        L.check(self._lib.aether_fluid_medium_phase_hg_pdf(self._h, _dptr(wo), _dptr(wi),
                                                           len(wo), _dptr(out)))
        return out

    def sample_phase_hg(self, wo, states):
        """src/fluid.rs:892-910 for every row of wo; returns (wi, pdf)."""
        wo = np.ascontiguousarray(np.asarray(wo, dtype=np.float64).reshape(-1, 3))
        st = self._states(states)
        wi = np.empty_like(wo)
        pdf = np.empty(len(wo), dtype=np.float64)
        L.check(self._lib.aether_fluid_medium_sample_phase_hg(
            self._h, _dptr(wo), len(wo), st.ctypes.data_as(C.POINTER(C.c_uint64)), _dptr(wi),
            _dptr(pdf)))
        return wi, pdf

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.aether_fluid_medium_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
