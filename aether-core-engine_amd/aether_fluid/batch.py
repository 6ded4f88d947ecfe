"""Batched smoke grids over the C ABI: many independent FluidSolver3D
instances of different sizes advanced together, one launch per stage."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .solver import FluidTelemetryFrame, _dptr


class FluidBatch:
    """instances: list of dicts with keys nx, ny, nz and optionally h,
    vorticity_eps, buoyancy_beta, gravity_alpha, source_center,
    source_radius, source_density, source_temperature.  Shared knobs
    (solver_iters, cg_rel_tol, dissipation, cg_form, cg_z_cells) come from
    keyword arguments; cg_form CG_AUTO runs the four-wide CG when every
    nx % 4 == 0 (FP32), with cg_z_cells (2 or 4) z planes per thread, 4 when
    0 and not tuned, or the autotuner's choice; CG_FUSED runs the one-launch
    CG solved group by group (FP32, every nx % 4 == 0).  graph=True captures
    each step into a HIP graph and replays it."""

    def __init__(self, instances, precision: str = "fp32", solver_iters: int = 200,
                 cg_rel_tol: float = 1e-5, dissipation=(0.996, 0.996, 0.985),
                 cg_form: int = L.CG_AUTO, graph: bool = False, cg_z_cells: int = 0):
        self._lib = L.lib()
        n = len(instances)
        descs = (L.InstanceDesc * n)()
        for q, d in enumerate(instances):
            e = descs[q]
            e.nx, e.ny, e.nz = d["nx"], d["ny"], d["nz"]
            e.h = d.get("h", 0.0)
            e.vorticity_eps = d.get("vorticity_eps", 5.5)
            e.buoyancy_beta = d.get("buoyancy_beta", 6.8)
            e.gravity_alpha = d.get("gravity_alpha", 0.45)
            e.source_center = (C.c_double * 3)(*d.get("source_center", (0.0, 0.0, 0.0)))
            e.source_radius = d.get("source_radius", 0.0)
            e.source_density = d.get("source_density", 1.0)
            e.source_temperature = d.get("source_temperature", 1.0)
        p = L.Params()
        L.check(self._lib.aether_fluid_default_params(C.byref(p)))
        p.solver_kind, p.solver_iters, p.cg_rel_tol = 0, solver_iters, cg_rel_tol
        p.cg_form = cg_form
        p.cg_z_cells = cg_z_cells
        p.batch_graph = 1 if graph else 0
        (p.dissipation_velocity, p.dissipation_density,
         p.dissipation_temperature) = dissipation
        prec = {"fp32": L.FP32, "fp64": L.FP64}[precision]
        h = C.c_void_p()
        L.check(self._lib.aether_fluid_batch_create(descs, n, prec, C.byref(p), C.byref(h)))
        self._h = h
        self.instances = [dict(d) for d in instances]
        self.count = n

    def close(self):
        if getattr(self, "_h", None):
            self._lib.aether_fluid_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _n(self, q):
        d = self.instances[q]
        return d["nx"] * d["ny"] * d["nz"]

    def upload(self, q: int, name: str, values) -> None:
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        L.check(self._lib.aether_fluid_batch_upload(self._h, q, L.FIELDS[name], _dptr(a),
                                                    a.size))

    def download(self, q: int, name: str) -> np.ndarray:
        out = np.empty(self._n(q), dtype=np.float64)
        L.check(self._lib.aether_fluid_batch_download(self._h, q, L.FIELDS[name], _dptr(out),
                                                      out.size))
        return out

    def run(self, dt: float, times) -> list:
        """Returns frames[step][instance]."""
        ts = np.ascontiguousarray(np.asarray(times, dtype=np.float64))
        out = (L.Telemetry * max(len(ts) * self.count, 1))()
        L.check(self._lib.aether_fluid_batch_run(self._h, dt, _dptr(ts), len(ts), out))
        return [[FluidTelemetryFrame._from(out[s * self.count + q]) for q in range(self.count)]
                for s in range(len(ts))]

    def autotune_cg(self, iters: int = 20) -> dict:
        """Times the batch's CG in each candidate layout and tiling (four-wide
        with 2 and 4 z planes per thread, one cell per thread, one launch per
        iteration solved group by group) and keeps the fastest as the batch's
        choice for CG_AUTO and cg_z_cells 0 (pressure and divergence are
        restored).  iters > 0: that many iterations on synthetic right-hand
        sides; iters 0: whole solves of the last step's right-hand sides at
        the batch's tolerance and cap (after at least one run() step)."""
        us = (C.c_double * 4)()
        form, z = C.c_int(), C.c_int()
        L.check(self._lib.aether_fluid_batch_autotune_cg(self._h, int(iters), us,
                                                         C.byref(form), C.byref(z)))
        return {"choice": {L.CG_X4: "x4", L.CG_SCALAR: "scalar",
                           L.CG_FUSED: "fused"}[form.value],
                "z_cells": z.value,
                "us_per_iter": {k: round(us[a], 3)
                                for a, k in enumerate(("x4_z2", "x4_z4", "scalar_z8", "fused_z4"))
                                if us[a] >= 0}}

    def iterations(self) -> np.ndarray:
        """CG iterations of the last run, shape (steps, instances)."""
        cnt = C.c_int()
        L.check(self._lib.aether_fluid_batch_iterations(self._h, None, 0, C.byref(cnt)))
        buf = (C.c_int32 * max(cnt.value, 1))()
        L.check(self._lib.aether_fluid_batch_iterations(self._h, buf, cnt.value, C.byref(cnt)))
        return np.array(buf[:cnt.value]).reshape(-1, self.count)
