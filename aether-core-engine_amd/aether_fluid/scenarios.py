"""Smoke workloads used by the tests and the benchmark driver.

Each entry fixes the grid, time step, forcing, source and solver settings of
one workload; ``seeded_config_state`` draws its initial fields from the
reference's SplitMix64-seeded Xoshiro256++ stream (src/rng.rs), filling
arrays in ascending idx order: u, then v, then w over all cells, then
density over the central block.
"""
from __future__ import annotations

import numpy as np

from .solver import rng_fill

SMOKE_CONFIGS = {
    # 32x48x32, h = 1/32, zero initial state, CPU-sized
    "plain_32x48x32": dict(
        dims=(32, 48, 32), h=1.0 / 32, dt=0.02, steps=50, seed=None, rho_block=None,
        source=dict(center=(16.0, 4.0, 16.0), radius=4.0),
        alpha=0.05, beta=1.0, eps=0.3, cg_tol=1e-5, cg_max=200, sor_omega=1.7,
        rbgs_sweeps=40),
    # 128^3, h = 1/128, seeded divergent velocity, density block [32, 96)^3
    "plain_128": dict(
        dims=(128, 128, 128), h=1.0 / 128, dt=0.01, steps=100, seed=0xAE7E0001,
        rho_block=(32, 96), source=dict(center=(64.0, 10.0, 64.0), radius=12.0),
        alpha=0.05, beta=1.0, eps=0.3, cg_tol=1e-5, cg_max=200, sor_omega=1.8,
        rbgs_sweeps=60),
    # 256^3, h = 1/256, the mixed-precision projection workload
    "mixed_256": dict(
        dims=(256, 256, 256), h=1.0 / 256, dt=0.005, steps=20, seed=0xAE7E0002,
        rho_block=(64, 192), source=dict(center=(128.0, 20.0, 128.0), radius=24.0),
        alpha=0.05, beta=1.0, eps=0.3, cg_tol=1e-3, cg_max=200, sor_omega=1.8,
        rbgs_sweeps=60),
}


def seeded_config_state(name: str, dims=None, seed=None):
    """Initial (u, v, w, density) as flat f64 arrays for workload `name`.

    `dims`/`seed` override the workload's grid and seed (smaller test
    grids keep the same recipe)."""
    cfg = SMOKE_CONFIGS[name]
    nx, ny, nz = dims or cfg["dims"]
    n = nx * ny * nz
    seed = cfg["seed"] if seed is None else seed
    zeros = np.zeros(n)
    if seed is None:
        return {"u": zeros.copy(), "v": zeros.copy(), "w": zeros.copy(),
                "density": zeros.copy()}
    u, st = rng_fill(seed, -1.0, 1.0, n)
    v, st = rng_fill(st, -1.0, 1.0, n)
    w, st = rng_fill(st, -1.0, 1.0, n)
    rho = np.zeros(n)
    lo, hi = cfg["rho_block"]
    if dims is not None:  # scale the block to the overriding grid
        lo_f, hi_f = lo / cfg["dims"][0], hi / cfg["dims"][0]
        lo_i = [int(round(lo_f * d)) for d in (nx, ny, nz)]
        hi_i = [int(round(hi_f * d)) for d in (nx, ny, nz)]
    else:
        lo_i, hi_i = [lo] * 3, [hi] * 3
    cnt = (hi_i[0] - lo_i[0]) * (hi_i[1] - lo_i[1]) * (hi_i[2] - lo_i[2])
    vals, st = rng_fill(st, 0.0, 1.0, cnt)
    block = rho.reshape(nz, ny, nx)
    block[lo_i[2]:hi_i[2], lo_i[1]:hi_i[1], lo_i[0]:hi_i[0]] = vals.reshape(
        hi_i[2] - lo_i[2], hi_i[1] - lo_i[1], hi_i[0] - lo_i[0])
    return {"u": u, "v": v, "w": w, "density": rho}


def configure(solver, name: str, solver_kind: str = "cg") -> None:
    """Apply workload `name`'s forcing, source and solver settings."""
    cfg = SMOKE_CONFIGS[name]
    solver.gravity_alpha = cfg["alpha"]
    solver.buoyancy_beta = cfg["beta"]
    solver.vorticity_eps = cfg["eps"]
    solver.set_sphere_source(cfg["source"]["center"], cfg["source"]["radius"])
    if solver_kind == "cg":
        solver.solver_kind = 0
        solver.solver_iters = cfg["cg_max"]
        solver.cg_rel_tol = cfg["cg_tol"]
    else:
        solver.solver_kind = 1
        solver.solver_iters = cfg["rbgs_sweeps"]
        solver.sor_omega = cfg["sor_omega"]


BATCH_SIZES = ((32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 96, 128))


def batched_instances(count: int = 64, seed: int = 0xAE7E0004):
    """The batched workload: `count` grids with sizes drawn from BATCH_SIZES,
    a sphere source (radius max(2, nx/10)) at a uniform position in the
    bottom quarter, alpha ~ U[0.02, 0.1], eps ~ U[0.1, 0.5], beta = 1,
    h = 1/max(n), all drawn from the reference's seeded stream."""
    draws, _ = rng_fill(seed, 0.0, 1.0, 6 * count)
    out = []
    for q in range(count):
        u = draws[6 * q:6 * q + 6]
        nx, ny, nz = BATCH_SIZES[min(int(u[0] * len(BATCH_SIZES)), len(BATCH_SIZES) - 1)]
        r = max(2.0, nx / 10.0)
        cx = (r + 1) + u[1] * (nx - 2 * r - 3)
        cy = 1.0 + u[2] * (ny / 4.0 - 1.0)
        cz = (r + 1) + u[3] * (nz - 2 * r - 3)
        out.append(dict(nx=nx, ny=ny, nz=nz, h=0.0, vorticity_eps=0.1 + 0.4 * u[5],
                        buoyancy_beta=1.0, gravity_alpha=0.02 + 0.08 * u[4],
                        source_center=(cx, cy, cz), source_radius=r,
                        source_density=1.0, source_temperature=1.0))
    return out


def rollout_problem(n: int = 64, seed: int = 0xAE7E0003):
    """The differentiable-rollout workload scaled to an n^3 grid (64 in the
    spec): inflow patch i, k in [n/4, 3n/4), j in [1, 1 + n/16) with seeded
    U[0, 1] controls, Gaussian density target sigma = n/8 centred at
    (n/2, 3n/4, n/2), h = 1/n."""
    lo = (n // 4, 1, n // 4)
    hi = (3 * n // 4, 1 + max(1, n // 16), 3 * n // 4)
    cnt = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])
    controls, _ = rng_fill(seed, 0.0, 1.0, cnt)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sig = n / 8.0
    c = (n / 2.0, 3.0 * n / 4.0, n / 2.0)
    target = np.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) / (2 * sig * sig))
    return dict(patch_lo=lo, patch_hi=hi, controls=controls, target=target.ravel())
