"""Python mirror of the reference's ``FluidSolver3D`` (src/fluid.rs:23-673)
over the C ABI of include/aether_fluid.h.

Names, argument meaning and error behaviour follow the reference:
``FluidSolver3D(nx, ny, nz)`` raises for dimensions below 6 (the reference
asserts, src/fluid.rs:54); the pub knobs ``vorticity_eps``, ``buoyancy_beta``,
``gravity_alpha``, ``solver_kind`` and ``solver_iters`` are plain attributes
(src/fluid.rs:45-49) pushed to the device before the next call; the pub
arrays ``u, v, w, density, temperature, pressure, divergence`` are
properties that download (get) or upload (set) flat f64 arrays indexed by
``idx(i, j, k) = (k*ny + j)*nx + i``.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _lib as L


class PoissonSolverKind(enum.IntEnum):
    """src/fluid.rs:17-21"""
    ConjugateGradient = L.SOLVER_CG
    RedBlackGaussSeidel = L.SOLVER_RBGS


@dataclass
class FluidTelemetryFrame:
    """src/fluid.rs:6-15"""
    time_sec: float
    kinetic_energy: float
    enstrophy: float
    peak_vorticity: float
    total_mass: float
    center_of_mass: tuple

    @classmethod
    def _from(cls, t: L.Telemetry) -> "FluidTelemetryFrame":
        return cls(t.time_sec, t.kinetic_energy, t.enstrophy, t.peak_vorticity, t.total_mass,
                   tuple(t.center_of_mass))

    def as_array(self) -> np.ndarray:
        return np.array([self.time_sec, self.kinetic_energy, self.enstrophy,
                         self.peak_vorticity, self.total_mass, *self.center_of_mass])


_PARAM_NAMES = ("vorticity_eps", "buoyancy_beta", "gravity_alpha", "solver_kind",
                "solver_iters", "sor_omega", "cg_rel_tol", "dissipation_velocity",
                "dissipation_density", "dissipation_temperature", "source_kind",
                "source_center", "source_radius", "source_density", "source_temperature",
                "ir_inner_rel_tol", "ir_div_tol", "ir_max_outer", "cg_form", "cg_z_cells")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class FluidSolver3D:
    """Device-resident smoke solver; see the module docstring."""

    def __init__(self, nx: int, ny: int, nz: int, h: float = 0.0, precision: str = "fp32"):
        prec = {"fp32": L.FP32, "fp64": L.FP64, "fp32_ir64": L.FP32_IR64}[precision]
        self._lib = L.lib()
        handle = C.c_void_p()
        L.check(self._lib.aether_fluid_create(nx, ny, nz, float(h), prec, C.byref(handle)))
        self._h = handle
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        hh = C.c_double()
        L.check(self._lib.aether_fluid_dims(self._h, None, None, None, C.byref(hh), None))
        self.h = hh.value
        self.precision = precision
        p = L.Params()
        L.check(self._lib.aether_fluid_default_params(C.byref(p)))
        for name in _PARAM_NAMES:
            val = getattr(p, name)
            if name == "source_center":
                val = tuple(val)
            object.__setattr__(self, name, val)
        self._pushed = None

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.aether_fluid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------------- params
    def _params(self) -> L.Params:
        p = L.Params()
        for name in _PARAM_NAMES:
            val = getattr(self, name)
            if name == "source_center":
                p.source_center = (C.c_double * 3)(*[float(x) for x in val])
            elif name in ("solver_kind", "solver_iters", "source_kind", "ir_max_outer",
                          "cg_form", "cg_z_cells"):
                setattr(p, name, int(val))
            else:
                setattr(p, name, float(val))
        return p

    def _sync_params(self) -> None:
        key = tuple(getattr(self, n) for n in _PARAM_NAMES)
        if key != self._pushed:
            p = self._params()
            L.check(self._lib.aether_fluid_set_params(self._h, C.byref(p)))
            self._pushed = key

    def set_sphere_source(self, center, radius: float, density: float = 1.0,
                          temperature: float = 1.0) -> None:
        """Sphere source that sets density and temperature inside d^2 < r^2."""
        self.source_kind = L.SOURCE_SPHERE
        self.source_center = tuple(float(c) for c in center)
        self.source_radius = float(radius)
        self.source_density = float(density)
        self.source_temperature = float(temperature)

    # ----------------------------------------------------------------- fields
    def idx(self, i: int, j: int, k: int) -> int:
        """src/fluid.rs:88-90"""
        return (k * self.ny + j) * self.nx + i

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        L.check(self._lib.aether_fluid_download(self._h, L.FIELDS[name], _dptr(out), self.n))
        return out

    def upload(self, name: str, values) -> None:
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        L.check(self._lib.aether_fluid_upload(self._h, L.FIELDS[name], _dptr(a), a.size))

    def _field_prop(name):  # noqa: N805 - property factory
        return property(lambda self: self.download(name),
                        lambda self, val: self.upload(name, val))

    u = _field_prop("u")
    v = _field_prop("v")
    w = _field_prop("w")
    density = _field_prop("density")
    temperature = _field_prop("temperature")
    pressure = _field_prop("pressure")
    divergence = _field_prop("divergence")
    del _field_prop

    # ---------------------------------------------------------------- methods
    def step(self, sim_dt: float, timeline_sec: float) -> FluidTelemetryFrame:
        """src/fluid.rs:598-611"""
        self._sync_params()
        t = L.Telemetry()
        L.check(self._lib.aether_fluid_step(self._h, sim_dt, timeline_sec, C.byref(t)))
        return FluidTelemetryFrame._from(t)

    def run(self, sim_dt: float, times) -> list:
        """Many steps with one synchronisation (the src/main.rs:89-94 loop)."""
        self._sync_params()
        ts = np.ascontiguousarray(np.asarray(times, dtype=np.float64))
        out = (L.Telemetry * max(len(ts), 1))()
        L.check(self._lib.aether_fluid_run(self._h, sim_dt, _dptr(ts), len(ts), out))
        return [FluidTelemetryFrame._from(out[a]) for a in range(len(ts))]

    def advect_maccormack_in_place(self, field_kind: int, dt: float, dissipation: float):
        """src/fluid.rs:170-234"""
        L.check(self._lib.aether_fluid_advect_maccormack_in_place(self._h, field_kind, dt,
                                                                  dissipation))

    def enforce_wall_velocity_bc(self) -> None:
        """src/fluid.rs:236-253"""
        L.check(self._lib.aether_fluid_enforce_wall_velocity_bc(self._h))

    def apply_buoyancy_and_vorticity_confinement(self, dt: float) -> None:
        """src/fluid.rs:256-330"""
        self._sync_params()
        L.check(self._lib.aether_fluid_apply_buoyancy_and_vorticity_confinement(self._h, dt))

    def compute_divergence_norm(self) -> float:
        """src/fluid.rs:333-353"""
        out = C.c_double()
        L.check(self._lib.aether_fluid_compute_divergence_norm(self._h, C.byref(out)))
        return out.value

    def divergence_max(self) -> float:
        out = C.c_double()
        L.check(self._lib.aether_fluid_divergence_max(self._h, C.byref(out)))
        return out.value

    def project(self) -> None:
        """src/fluid.rs:496-529"""
        self._sync_params()
        L.check(self._lib.aether_fluid_project(self._h))

    def solve_poisson_conjugate_gradient(self, max_iters: int) -> None:
        """src/fluid.rs:445-493 (on the stored divergence and pressure)"""
        self._sync_params()
        L.check(self._lib.aether_fluid_solve_poisson_conjugate_gradient(self._h, max_iters))

    def solve_poisson_gauss_seidel(self, iterations: int) -> None:
        """src/fluid.rs:397-442"""
        self._sync_params()
        L.check(self._lib.aether_fluid_solve_poisson_gauss_seidel(self._h, iterations))

    def inject_sources(self, time_sec: float, dt: float) -> None:
        self._sync_params()
        L.check(self._lib.aether_fluid_inject_sources(self._h, time_sec, dt))

    def inject_cyberpunk_sources(self, time_sec: float, dt: float) -> None:
        """src/fluid.rs:532-595 (the reference's jets, whatever source_kind says)"""
        saved = self.source_kind
        self.source_kind = L.SOURCE_JETS
        try:
            self.inject_sources(time_sec, dt)
        finally:
            self.source_kind = saved

    def compute_telemetry(self, timeline_sec: float) -> FluidTelemetryFrame:
        """src/fluid.rs:613-672"""
        t = L.Telemetry()
        L.check(self._lib.aether_fluid_compute_telemetry(self._h, timeline_sec, C.byref(t)))
        return FluidTelemetryFrame._from(t)

    def sample_trilinear(self, field, x, y=None, z=None):
        """src/fluid.rs:94-128; scalar (x, y, z) or an (n, 3) array of points.
        `field` is the name of one of the solver's fields, or (as the
        reference's signature takes) any f64 array of nx*ny*nz values."""
        pts = np.asarray([[x, y, z]] if y is not None else x, dtype=np.float64).reshape(-1, 3)
        pts = np.ascontiguousarray(pts)
        out = np.empty(len(pts), dtype=np.float64)
        if isinstance(field, str):
            L.check(self._lib.aether_fluid_sample_trilinear(self._h, L.FIELDS[field], _dptr(pts),
                                                            len(pts), _dptr(out)))
        else:
            f = np.ascontiguousarray(np.asarray(field, dtype=np.float64).reshape(-1))
            L.check(self._lib.aether_fluid_sample_trilinear_field(self._h, _dptr(f), f.size,
                                                                  _dptr(pts), len(pts),
                                                                  _dptr(out)))
        return float(out[0]) if y is not None else out

    def nonfinite(self) -> bool:
        """True if a telemetry reduction since the last call met a non-finite
        velocity, vorticity or density (the flag is then cleared)."""
        f = C.c_int()
        L.check(self._lib.aether_fluid_nonfinite(self._h, C.byref(f)))
        return bool(f.value)

    def max_density(self) -> float:
        out = C.c_double()
        L.check(self._lib.aether_fluid_max_density(self._h, C.byref(out)))
        return out.value

    def last_solve_stats(self) -> dict:
        s = L.SolveStats()
        L.check(self._lib.aether_fluid_last_solve_stats(self._h, C.byref(s)))
        return {"iters": s.iterations, "converged": bool(s.converged), "b2": s.b_norm2,
                "r2_init": s.r_norm2_initial, "r2": s.r_norm2_final,
                "outer": s.outer_iterations, "div_max": s.div_max_final}

    def run_iters(self) -> list:
        """Per-step CG iterations (RBGS: sweeps) of the last run/step."""
        cnt = C.c_int()
        L.check(self._lib.aether_fluid_last_run_iterations(self._h, None, 0, C.byref(cnt)))
        buf = (C.c_int32 * max(cnt.value, 1))()
        L.check(self._lib.aether_fluid_last_run_iterations(self._h, buf, cnt.value,
                                                           C.byref(cnt)))
        return [buf[a] for a in range(cnt.value)]

    def rollout(self, nsteps: int, dt: float, patch_lo, patch_hi, controls, beta: float,
                target, checkpoint_every: int = 8, grad: bool = True,
                adjoint_cg_iters: int | None = None, adjoint_cg_rel_tol: float | None = None):
        """Differentiable rollout from the current state (left unchanged).

        controls: inflow density/temperature of the patch [patch_lo, patch_hi)
        (x fastest), beta: buoyancy coefficient, target: density target.
        Returns (loss, dloss/dcontrols, dloss/dbeta); the gradients are None
        when grad is False."""
        self._sync_params()
        d = self._rollout_desc(nsteps, dt, patch_lo, patch_hi, checkpoint_every,
                               adjoint_cg_iters, adjoint_cg_rel_tol)
        c = np.ascontiguousarray(np.asarray(controls, dtype=np.float64).reshape(-1))
        t = np.ascontiguousarray(np.asarray(target, dtype=np.float64).reshape(-1))
        loss = C.c_double()
        gc = np.zeros_like(c) if grad else None
        gb = C.c_double()
        L.check(self._lib.aether_fluid_rollout(
            self._h, C.byref(d), _dptr(c), float(beta), _dptr(t), C.byref(loss),
            _dptr(gc) if grad else None, C.byref(gb) if grad else None))
        return loss.value, gc, (gb.value if grad else None)

    STATE_FIELDS = ("u", "v", "w", "density", "temperature")

    def _rollout_desc(self, nsteps, dt, patch_lo, patch_hi, checkpoint_every,
                      adjoint_cg_iters=None, adjoint_cg_rel_tol=None):
        d = L.RolloutDesc()
        d.nsteps, d.checkpoint_every, d.dt = nsteps, checkpoint_every, dt
        d.patch_lo = (C.c_int32 * 3)(*patch_lo)
        d.patch_hi = (C.c_int32 * 3)(*patch_hi)
        d.adjoint_cg_iters = self.solver_iters if adjoint_cg_iters is None else adjoint_cg_iters
        d.adjoint_cg_rel_tol = (self.cg_rel_tol if adjoint_cg_rel_tol is None
                                else adjoint_cg_rel_tol)
        return d

    def rollout_forward(self, nsteps: int, dt: float, patch_lo, patch_hi, controls, beta: float,
                        checkpoint_every: int = 8, adjoint_cg_iters: int | None = None,
                        adjoint_cg_rel_tol: float | None = None) -> dict:
        """Forward sweep of a differentiable rollout from the current state
        (left unchanged), recording the tape for rollout_backward.  Returns
        the final state {u, v, w, density, temperature} (flat f64)."""
        self._sync_params()
        d = self._rollout_desc(nsteps, dt, patch_lo, patch_hi, checkpoint_every,
                               adjoint_cg_iters, adjoint_cg_rel_tol)
        c = np.ascontiguousarray(np.asarray(controls, dtype=np.float64).reshape(-1))
        out = np.empty(5 * self.n, dtype=np.float64)
        L.check(self._lib.aether_fluid_rollout_forward(self._h, C.byref(d), _dptr(c),
                                                       float(beta), _dptr(out)))
        self._tape_nctrl = c.size
        return {name: out[a * self.n:(a + 1) * self.n] for a, name in enumerate(self.STATE_FIELDS)}

    def rollout_backward(self, lam_final: dict | None):
        """Vector-Jacobian product of the recorded rollout: lam_final holds
        the adjoints of the final fields (missing names are zero).  Returns
        (d/dcontrols, d/dbeta, {field: d/d initial field})."""
        lam = None
        if lam_final is not None:
            lam = np.zeros(5 * self.n, dtype=np.float64)
            for a, name in enumerate(self.STATE_FIELDS):
                if name in lam_final and lam_final[name] is not None:
                    lam[a * self.n:(a + 1) * self.n] = np.asarray(lam_final[name],
                                                                  dtype=np.float64).reshape(-1)
        gc = np.zeros(getattr(self, "_tape_nctrl", 0), dtype=np.float64)
        gb = C.c_double()
        gi = np.empty(5 * self.n, dtype=np.float64)
        L.check(self._lib.aether_fluid_rollout_backward(
            self._h, _dptr(lam) if lam is not None else None, _dptr(gc), C.byref(gb), _dptr(gi)))
        return gc, gb.value, {name: gi[a * self.n:(a + 1) * self.n]
                              for a, name in enumerate(self.STATE_FIELDS)}

    _TUNE_CANDIDATES = (("x4", 2), ("x4", 4), ("scalar", 4), ("scalar", 8), ("scalar", 16),
                        ("fused", 4))

    def autotune_cg(self, iters: int = 30) -> dict:
        """Times `iters` CG iterations of every layout candidate on a synthetic
        right-hand side and keeps the fastest as this solver's CG_AUTO choice
        (pressure and divergence are restored).  Returns the choice and the
        time per iteration (us) of each candidate that applies."""
        us = (C.c_double * 6)()
        form, z = C.c_int(), C.c_int()
        L.check(self._lib.aether_fluid_autotune_cg(self._h, int(iters), us, C.byref(form),
                                                   C.byref(z)))
        names = {L.CG_SCALAR: "scalar", L.CG_X4: "x4", L.CG_FUSED: "fused"}
        return {"choice": {"form": names[form.value], "z_cells": z.value},
                "us_per_iter": {f"{f}_z{zc}": round(us[a], 3)
                                for a, (f, zc) in enumerate(self._TUNE_CANDIDATES) if us[a] >= 0}}

    def set_kernel_timing(self, enable: bool) -> None:
        L.check(self._lib.aether_fluid_set_kernel_timing(self._h, int(enable)))

    def kernel_timing(self, which: int) -> tuple:
        avg = C.c_double()
        cnt = C.c_long()
        L.check(self._lib.aether_fluid_kernel_timing(self._h, which, C.byref(avg),
                                                     C.byref(cnt)))
        return avg.value, cnt.value


def rng_fill(seed_or_state, lo: float, hi: float, n: int):
    """Draw n range_f64(lo, hi) values from the reference's seeded stream
    (src/rng.rs:21-33, 86-88).  Returns (values, state) so that draws can
    continue from the same stream."""
    lib = L.lib()
    st = (C.c_uint64 * 4)()
    if isinstance(seed_or_state, int):
        L.check(lib.aether_rng_seed(C.c_uint64(seed_or_state), st))
    else:
        for a in range(4):
            st[a] = seed_or_state[a]
    out = np.empty(n, dtype=np.float64)
    L.check(lib.aether_rng_fill_range(st, lo, hi, _dptr(out), n))
    return out, [st[a] for a in range(4)]


def stream_copy_ms(nbytes: int, reps: int = 20) -> float:
    """Average device time (ms) of a plain nbytes copy kernel."""
    out = C.c_double()
    L.check(L.lib().aether_fluid_stream_copy(nbytes, reps, C.byref(out)))
    return out.value
