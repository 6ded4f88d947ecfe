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


def advection_switch_cells(u, v, w, dims, scale, delta=1e-6):
    """Cells whose MacCormack result (src/fluid.rs:170-234) can jump when a
    trace coordinate moves by `delta`, found from the (shared) velocity:

    * a back-trace i - scale*u (and j, k), u != 0, whose clamped value lies
      within delta of an integer: floor() picks the other corner cell for
      the forward sample and the 8-corner extrema (fluid.rs:99-101,
      137-139);
    * a back-trace within delta of the near-wall thresholds 1.5 and
      n - 2.5, which switch between the corrected and the forward value
      (fluid.rs:196-201);
    * a forward trace i + scale*u, u != 0 (the backward semi-Lagrangian
      pass, fluid.rs:184), within delta of an integer;
    * a cell whose forward trace has, among its 8 corners, a cell with a
      switching back-trace: the backward pass reads that cell's forward
      value.

    Returns (mask of all of these, shaped (nz, ny, nx), and the number of
    primary switch cells, the first three kinds).  Boundary cells are never
    traced."""
    nx, ny, nz = dims
    shape = (nz, ny, nx)
    u, v, w = (np.asarray(a, dtype=np.float64).reshape(shape) for a in (u, v, w))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    interior = ((i >= 1) & (i <= nx - 2) & (j >= 1) & (j <= ny - 2) & (k >= 1) & (k <= nz - 2))
    back = np.zeros(shape, dtype=bool)
    fwd = np.zeros(shape, dtype=bool)
    corner0 = []
    for idx, vel, n in ((i, u, nx), (j, v, ny), (k, w, nz)):
        for sign, out in ((-1.0, back), (1.0, fwd)):
            x = idx + sign * scale * vel
            xc = np.clip(x, 0.5, n - 1.500001)
            # a zero velocity component traces to i itself in any precision:
            # no switch there (the wall masks leave many exact zeros)
            out |= (np.abs(xc - np.round(xc)) < delta) & (vel != 0.0)
            if sign < 0:
                out |= (np.abs(x - 1.5) < delta) | (np.abs(x - (n - 2.5)) < delta)
            else:
                corner0.append(np.floor(xc).astype(np.int64))
    back &= interior
    fwd &= interior
    i0, j0, k0 = corner0
    reads = np.zeros(shape, dtype=bool)
    if back.any():
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    reads |= back[k0 + dk, j0 + dj, i0 + di]
        reads &= interior
    return back | fwd | reads, int((back | fwd).sum())
