"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is the f64 C restatement of the reference `FluidSolver3D`
(src/fluid.rs:23-673) in ``oracle/fluid_oracle.c``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product package never does.

``OracleFluid`` mirrors the reference's public surface: the constructor, the
parameter fields, ``step``, ``project``, ``compute_divergence_norm``,
``inject_sources``, ``apply_buoyancy_and_vorticity_confinement``,
``enforce_wall_velocity_bc``, ``advect_maccormack_in_place``,
``compute_telemetry`` and ``sample_trilinear``; fields are exposed as numpy
views shaped (nz, ny, nx) over the oracle's own buffers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = {"u": 0, "v": 1, "w": 2, "density": 3, "temperature": 4, "pressure": 5,
          "divergence": 6, "curl_mag": 7, "cg_r": 8}

_libs: dict = {}


def build() -> None:
    """Compile both oracle builds with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(omp: bool = False) -> C.CDLL:
    name = "liboracle_fluid_omp.so" if omp else "liboracle_fluid.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    vp, d, l, i = C.c_void_p, C.c_double, C.c_long, C.c_int
    pd = C.POINTER(C.c_double)
    sig = {
        "of_create": (vp, [l, l, l, d]),
        "of_destroy": (None, [vp]),
        "of_set_params": (None, [vp, d, d, d, i, l, d, d, d, d, d]),
        "of_set_source": (None, [vp, i, d, d, d, d, d, d]),
        "of_field": (pd, [vp, i]),
        "of_h": (d, [vp]),
        "of_stats": (None, [vp, C.POINTER(l), pd, pd, pd, C.POINTER(i)]),
        "of_sample_trilinear": (d, [vp, i, d, d, d]),
        "of_max_density": (d, [vp]),
        "of_soundtrack": (None, [pd, l, C.POINTER(C.c_int32)]),
        "of_soundtrack_ex": (None, [pd, l, d, d, l, C.POINTER(C.c_int32)]),
        "of_medium_sample": (None, [vp, pd, pd, d, d, d, pd, pd]),
        "of_advect": (None, [vp, i, d, d]),
        "of_enforce_bc": (None, [vp]),
        "of_buoyancy_confinement": (None, [vp, d]),
        "of_divergence_norm": (d, [vp]),
        "of_divergence_max": (d, [vp]),
        "of_apply_laplacian": (None, [vp, pd, pd]),
        "of_solve_rbgs": (None, [vp, l]),
        "of_solve_cg": (None, [vp, l]),
        "of_project": (None, [vp]),
        "of_inject_sources": (None, [vp, d, d]),
        "of_telemetry": (None, [vp, d, pd]),
        "of_step": (None, [vp, d, d, pd]),
        "of_num_threads": (i, []),
        "orng_seed": (None, [vp, C.c_uint64]),
        "orng_set_state": (None, [vp, C.POINTER(C.c_uint64)]),
        "orng_next_u64": (C.c_uint64, [vp]),
        "orng_next_f64": (d, [vp]),
        "orng_range_f64": (d, [vp, d, d]),
        "of_splitmix64_next": (C.c_uint64, [C.POINTER(C.c_uint64)]),
        "of_medium_interaction_batch": (None, [vp, vp, l, pd, pd, pd, pd, pd, vp, pd]),
        "of_medium_transmittance_batch": (None, [vp, vp, l, pd, pd, pd, pd, vp, pd]),
        "of_phase_sample_batch": (None, [d, l, pd, vp, pd, pd]),
        "of_phase_pdf_batch": (None, [d, l, pd, pd, pd]),
        "of_normalize_batch": (None, [l, pd, pd]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def soundtrack(frames, max_ke: float = 0.0, max_vort: float = 0.0,
               nsamples: int = 720000) -> np.ndarray:
    """audio.rs:397-604: the 15 s, 48 kHz stereo soundtrack of a telemetry
    sequence (rows of 8: time, KE, enstrophy, peak |w|, mass, COM xyz),
    quantised to 24-bit PCM with the TPDF dither; shape (nsamples, 2).
    max_ke / max_vort > 0 replace the normalisers (the maxima over the
    frames)."""
    lib = _load()
    fr = np.ascontiguousarray(np.asarray(frames, dtype=np.float64).reshape(-1, 8))
    out = np.empty(2 * nsamples, dtype=np.int32)
    lib.of_soundtrack_ex(fr.ctypes.data_as(C.POINTER(C.c_double)), len(fr), float(max_ke),
                         float(max_vort), nsamples, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(-1, 2)


class OracleRng:
    """SplitMix64-seeded Xoshiro256++ (src/rng.rs:5-88)."""

    def __init__(self, seed: int | None = None, state=None):
        self._lib = _load()
        self._buf = C.create_string_buffer(32)
        if state is not None:
            arr = (C.c_uint64 * 4)(*state)
            self._lib.orng_set_state(self._buf, arr)
        else:
            self._lib.orng_seed(self._buf, C.c_uint64(seed))

    def next_u64(self) -> int:
        return self._lib.orng_next_u64(self._buf)

    def next_f64(self) -> float:
        return self._lib.orng_next_f64(self._buf)

    def range_f64(self, lo: float, hi: float) -> float:
        return self._lib.orng_range_f64(self._buf, lo, hi)


def splitmix64_stream(seed: int, n: int) -> list:
    lib = _load()
    st = C.c_uint64(seed)
    return [lib.of_splitmix64_next(C.byref(st)) for _ in range(n)]


class OracleFluid:
    """The f64 oracle solver; constructor and fields follow src/fluid.rs:23-85."""

    def __init__(self, nx: int, ny: int, nz: int, h: float = 0.0, omp: bool = False):
        self._lib = _load(omp)
        self._h = self._lib.of_create(nx, ny, nz, h)
        if not self._h:
            raise ValueError("grid dimensions must be >= 6 (src/fluid.rs:54)")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        self.h = self._lib.of_h(self._h)
        self.vorticity_eps = 5.5
        self.buoyancy_beta = 6.8
        self.gravity_alpha = 0.45
        self.solver_kind = 0
        self.solver_iters = 32
        self.sor_omega = 1.65
        self.cg_rel_tol = 0.0
        self.dissipation = (0.996, 0.996, 0.985)
        self._push()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.of_destroy(self._h)
            self._h = None

    def _push(self):
        dv, dr, dt = self.dissipation
        self._lib.of_set_params(self._h, self.vorticity_eps, self.buoyancy_beta,
                                self.gravity_alpha, int(self.solver_kind),
                                int(self.solver_iters), self.sor_omega, self.cg_rel_tol,
                                dv, dr, dt)

    def set_source(self, kind: int, center=(0.0, 0.0, 0.0), radius=0.0, rho=1.0, temp=1.0):
        self._lib.of_set_source(self._h, kind, float(center[0]), float(center[1]),
                                float(center[2]), float(radius), float(rho), float(temp))

    def field(self, name: str) -> np.ndarray:
        p = self._lib.of_field(self._h, FIELDS[name])
        return np.ctypeslib.as_array(p, shape=(self.nz, self.ny, self.nx))

    def __getattr__(self, name):
        if name in FIELDS:
            return self.field(name)
        raise AttributeError(name)

    # --- reference methods -------------------------------------------------
    def step(self, dt: float, t: float) -> np.ndarray:
        self._push()
        out = np.zeros(8)
        self._lib.of_step(self._h, dt, t, _ptr(out))
        return out

    def advect_maccormack_in_place(self, kind: int, dt: float, dissipation: float):
        self._lib.of_advect(self._h, kind, dt, dissipation)

    def enforce_wall_velocity_bc(self):
        self._lib.of_enforce_bc(self._h)

    def apply_buoyancy_and_vorticity_confinement(self, dt: float):
        self._push()
        self._lib.of_buoyancy_confinement(self._h, dt)

    def compute_divergence_norm(self) -> float:
        return self._lib.of_divergence_norm(self._h)

    def divergence_max(self) -> float:
        return self._lib.of_divergence_max(self._h)

    def project(self):
        self._push()
        self._lib.of_project(self._h)

    def solve_poisson_conjugate_gradient(self, iters: int):
        self._push()
        self._lib.of_solve_cg(self._h, iters)

    def solve_poisson_gauss_seidel(self, iters: int):
        self._push()
        self._lib.of_solve_rbgs(self._h, iters)

    def inject_sources(self, t: float, dt: float):
        self._lib.of_inject_sources(self._h, t, dt)

    def compute_telemetry(self, t: float) -> np.ndarray:
        out = np.zeros(8)
        self._lib.of_telemetry(self._h, t, _ptr(out))
        return out

    def sample_trilinear(self, field: str, x: float, y: float, z: float) -> float:
        return self._lib.of_sample_trilinear(self._h, FIELDS[field], x, y, z)

    def max_density(self) -> float:
        """fluid.rs:704-708 (FluidMedium::from_solver's max over density)"""
        return self._lib.of_max_density(self._h)

    def medium_sample(self, bmin, bmax, points) -> np.ndarray:
        """fluid.rs:739-788 FluidMedium::sample_density_and_temp at each
        point (rows of xyz); returns an (n, 2) array of (density, temp)."""
        lo = (C.c_double * 3)(*bmin)
        hi = (C.c_double * 3)(*bmax)
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.empty((len(pts), 2))
        dv, tv = C.c_double(), C.c_double()
        for a, (x, y, z) in enumerate(pts):
            self._lib.of_medium_sample(self._h, lo, hi, float(x), float(y), float(z),
                                       C.byref(dv), C.byref(tv))
            out[a] = (dv.value, tv.value)
        return out

    # --- FluidMedium volumetric sampling (fluid.rs:793-910) ---------------
    def _omedium(self, bmin, bmax, sigma_t_scale, phase_g):
        """FluidMedium::from_solver's parameters (fluid.rs:703-735) over the
        oracle's current density and temperature."""
        md = self.max_density()
        return (C.c_double * 10)(*bmin, *bmax, sigma_t_scale, md,
                                 max(md * sigma_t_scale, 1e-4), min(max(phase_g, -0.9), 0.9))

    def medium_interaction(self, bmin, bmax, sigma_t_scale, phase_g, origins, dirs, t_min,
                           t_far, lambdas, states):
        """sample_interaction for each ray (dirs already normalised, as
        Ray::dir); states (n, 4) uint64 are advanced in place.  Returns an
        (n, 14) array: hit, t, point xyz, albedo x4, emission x4."""
        m = self._omedium(bmin, bmax, sigma_t_scale, phase_g)
        o, dd, t0, t1, lam = (np.ascontiguousarray(a, dtype=np.float64)
                              for a in (origins, dirs, t_min, t_far, lambdas))
        assert states.dtype == np.uint64 and states.flags.c_contiguous
        out = np.empty((len(t0), 14))
        self._lib.of_medium_interaction_batch(self._h, C.cast(m, C.c_void_p), len(t0), _ptr(o),
                                              _ptr(dd), _ptr(t0), _ptr(t1), _ptr(lam),
                                              states.ctypes.data, _ptr(out))
        return out

    def medium_transmittance(self, bmin, bmax, sigma_t_scale, phase_g, origins, dirs, t_min,
                             t_far, states):
        m = self._omedium(bmin, bmax, sigma_t_scale, phase_g)
        o, dd, t0, t1 = (np.ascontiguousarray(a, dtype=np.float64)
                         for a in (origins, dirs, t_min, t_far))
        assert states.dtype == np.uint64 and states.flags.c_contiguous
        out = np.empty(len(t0))
        self._lib.of_medium_transmittance_batch(self._h, C.cast(m, C.c_void_p), len(t0), _ptr(o),
                                                _ptr(dd), _ptr(t0), _ptr(t1), states.ctypes.data,
                                                _ptr(out))
        return out

    def apply_laplacian(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.nz, self.ny, self.nx)
        out = np.zeros_like(x)
        self._lib.of_apply_laplacian(self._h, _ptr(x), _ptr(out))
        return out

    def stats(self) -> dict:
        it = C.c_long()
        b2, r0, r2 = C.c_double(), C.c_double(), C.c_double()
        conv = C.c_int()
        self._lib.of_stats(self._h, C.byref(it), C.byref(b2), C.byref(r0), C.byref(r2),
                           C.byref(conv))
        return {"iters": it.value, "b2": b2.value, "r2_init": r0.value, "r2": r2.value,
                "converged": bool(conv.value)}

    @staticmethod
    def phase_sample(g, wo, states):
        """sample_phase_hg (fluid.rs:892-910) for each row of wo with phase_g
        g (already clamped); returns (wi (n, 3), pdf (n))."""
        lib = _load(False)
        wo = np.ascontiguousarray(wo, dtype=np.float64)
        wi, pdf = np.empty_like(wo), np.empty(len(wo))
        assert states.dtype == np.uint64 and states.flags.c_contiguous
        lib.of_phase_sample_batch(float(g), len(wo), _ptr(wo), states.ctypes.data, _ptr(wi),
                                  _ptr(pdf))
        return wi, pdf

    @staticmethod
    def phase_pdf(g, wo, wi):
        lib = _load(False)
        wo = np.ascontiguousarray(wo, dtype=np.float64)
        wi = np.ascontiguousarray(wi, dtype=np.float64)
        pdf = np.empty(len(wo))
        lib.of_phase_pdf_batch(float(g), len(wo), _ptr(wo), _ptr(wi), _ptr(pdf))
        return pdf

    @staticmethod
    def normalize(v):
        """Vec3::normalize (math.rs:58-65) of each row: Ray::new's dir."""
        lib = _load(False)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty_like(v)
        lib.of_normalize_batch(len(v), _ptr(v), _ptr(out))
        return out

    @staticmethod
    def num_threads(omp: bool = False) -> int:
        return _load(omp).of_num_threads()
