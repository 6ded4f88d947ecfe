"""Shared set-up for parity tests: one HIP solver and one oracle solver with
identical state and knobs."""
from __future__ import annotations

import numpy as np

FIELD_NAMES = ("u", "v", "w", "density", "temperature", "pressure")


def random_state(nx, ny, nz, seed=1, vel=1.0):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    return {
        "u": rng.uniform(-vel, vel, n),
        "v": rng.uniform(-vel, vel, n),
        "w": rng.uniform(-vel, vel, n),
        "density": rng.uniform(0.0, 1.0, n),
        "temperature": rng.uniform(0.0, 1.0, n),
        "pressure": rng.uniform(-0.05, 0.05, n),
    }


def load_state(gpu, orc, state):
    for name, arr in state.items():
        gpu.upload(name, arr)
        orc.field(name)[...] = np.asarray(arr).reshape(orc.nz, orc.ny, orc.nx)


def make_pair(nx, ny, nz, h=0.0, precision="fp64", state=None):
    import aether_fluid as af
    from oracle.oracle import OracleFluid
    gpu = af.FluidSolver3D(nx, ny, nz, h=h, precision=precision)
    orc = OracleFluid(nx, ny, nz, h=h)
    if state is not None:
        load_state(gpu, orc, state)
    return gpu, orc


def set_knobs(gpu, orc, *, eps=5.5, beta=6.8, alpha=0.45, kind=0, iters=32, omega=1.65,
              rel_tol=0.0, diss=(0.996, 0.996, 0.985), source=None):
    """source: None (reference jets) or dict(center=(x,y,z), radius=r)."""
    for s in (gpu, orc):
        s.vorticity_eps, s.buoyancy_beta, s.gravity_alpha = eps, beta, alpha
        s.solver_kind, s.solver_iters = kind, iters
        s.sor_omega, s.cg_rel_tol = omega, rel_tol
    gpu.dissipation_velocity, gpu.dissipation_density, gpu.dissipation_temperature = diss
    orc.dissipation = diss
    if source is None:
        gpu.source_kind = 0
        orc.set_source(0)
    else:
        gpu.set_sphere_source(source["center"], source["radius"])
        orc.set_source(1, source["center"], source["radius"], 1.0, 1.0)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def compare_fields(gpu, orc, names=FIELD_NAMES):
    return {n: rel_l2(gpu.download(n), orc.field(n)) for n in names}
