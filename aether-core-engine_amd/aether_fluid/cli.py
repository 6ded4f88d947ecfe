"""Command line of the smoke solver: the reference's fluid stage
(src/main.rs:82-113) and its projection self-check (src/verify.rs:229-250),
run on the GPU through the C ABI.

    python -m aether_fluid selfcheck [--precision fp64|fp32]
    python -m aether_fluid run [--steps 120] [--precision fp64|fp32]
"""
from __future__ import annotations

import argparse
import sys
import time

from . import _lib as L
from .medium import FluidMedium
from .solver import FluidSolver3D

# README.md:54: the reference's printed result of the self-check scenario
KNOWN_ANSWER = ("2.5751e+00", "6.7289e-03", "99.7")


def selfcheck(precision: str = "fp64") -> dict:
    """src/verify.rs:229-250: 16x20x16, CG 40 iterations, jets at t = 0.5,
    confinement/buoyancy 0.1, then one projection; passes when the RMS
    divergence drops to <= 3% of its value."""
    sim = FluidSolver3D(16, 20, 16, precision=precision)
    sim.solver_kind, sim.solver_iters = L.SOLVER_CG, 40
    sim.inject_cyberpunk_sources(0.5, 0.1)
    sim.apply_buoyancy_and_vorticity_confinement(0.1)
    pre = sim.compute_divergence_norm()
    sim.project()
    post = sim.compute_divergence_norm()
    digits = (f"{pre:.4e}", f"{post:.4e}", f"{(1.0 - post / pre) * 100.0:.1f}")
    return {"name": "3D Navier-Stokes CG Poisson Projection", "passed": post <= pre * 0.03,
            "detail": f"||div u||_2: {digits[0]} -> {digits[1]} ({digits[2]}% reduction)",
            "pre": pre, "post": post, "matches_reference_digits": digits == KNOWN_ANSWER}


def fluid_stage(steps: int = 120, precision: str = "fp64") -> dict:
    """src/main.rs:82-113: 36x52x36, vorticity_eps 6.5, 28 CG iterations,
    dt 0.08, timeline 15 s * step / (steps - 1), the reference's jets; then
    FluidMedium::from_solver over the scene's medium bounds."""
    nx, ny, nz = 36, 52, 36
    solver = FluidSolver3D(nx, ny, nz, precision=precision)
    solver.vorticity_eps = 6.5
    solver.solver_iters = 28
    times = [15.0 * (i / max(steps - 1, 1)) for i in range(steps)]
    t0 = time.perf_counter()
    frames = solver.run(0.08, times)
    elapsed = time.perf_counter() - t0
    peak_ke = max([0.0] + [f.kinetic_energy for f in frames])
    peak_vort = max([0.0] + [f.peak_vorticity for f in frames])
    medium = FluidMedium.from_solver(solver, ((-2.25, 0.22, -2.15), (2.25, 5.15, 1.45)), 1.42, 0.34)
    line = (f" [1/4] 3D Navier-Stokes Solver ({nx}x{ny}x{nz}, {steps} steps): {elapsed:.3f}s | "
            f"Peak KE={peak_ke:.4e} | Peak |w|={peak_vort:.2f} | Max rho={medium.max_density:.2f}")
    return {"line": line, "frames": frames, "peak_ke": peak_ke, "peak_vort": peak_vort,
            "max_density": medium.max_density, "seconds": elapsed}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m aether_fluid")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("selfcheck", help="src/verify.rs:229-250 on the GPU")
    a.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    b = sub.add_parser("run", help="the src/main.rs:82-113 fluid stage on the GPU")
    b.add_argument("--steps", type=int, default=120)
    b.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    args = ap.parse_args(argv)
    if args.cmd == "selfcheck":
        r = selfcheck(args.precision)
        print(f"[{'PASS' if r['passed'] else 'FAIL'}] {r['name']}: {r['detail']}"
              f"{'' if r['matches_reference_digits'] else ' (digits differ from the reference)'}")
        return 0 if r["passed"] else 1
    r = fluid_stage(max(args.steps, 10), args.precision)  # main.rs:39 clamps to >= 10
    print(r["line"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
