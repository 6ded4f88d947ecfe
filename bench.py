"""Benchmark: FP32 smoke step at 128^3 on one MI355X.

Workload: 128^3 MAC grid, h = 1/128, dt = 0.01, seeded strongly divergent
initial velocity (u, v, w i.i.d. uniform[-1, 1]) and density block [32, 96)^3,
sphere source r = 12 at (64, 10, 64), Boussinesq alpha = 0.05, beta = 1,
confinement eps = 0.3, MacCormack with the 8-cell clamp, CG projection to
relative residual 1e-5 (max 200 iterations).  A "step" is one full
FluidSolver3D::step (src/fluid.rs:598-611) of that grid.

Prints one JSON line: voxel-steps/s (value), the HBM roofline of the
dominant kernel (k_cg_cgs4, the one-launch CG iteration) measured with HIP
events on the solver's stream, and the CPU oracle timed on this host's cores
(cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "aether-core-engine_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = ("smoke voxel-steps/s at 128³ (fwd, fwd+adjoint), % HBM roofline, vs host CPU")
HBM_PEAK_GBPS = 8000.0
WORKLOAD = "plain_128"
SMOKE_DT = 0.01
COPY_BYTES = 256 << 20


def algorithmic_bytes_per_step(nvox: int, cg_iters: int) -> float:
    """FP32 array passes of one step with k CG iterations: 4 B * N * (59 + 8k)."""
    return 4.0 * nvox * (59 + 8 * cg_iters)


def algorithmic_bytes_per_rbgs_step(nvox: int, sweeps: int) -> float:
    """FP32 array passes of one red-black SOR step with s sweeps: the stages
    with the divergence writing b only (58), and per sweep two colour passes
    that each read p and their half of b and write their half of p (4):
    4 B * N * (58 + 4s)."""
    return 4.0 * nvox * (58 + 4 * sweeps)


def make_solver(solver_kind: str):
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, configure, seeded_config_state
    cfg = SMOKE_CONFIGS[WORKLOAD]
    s = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision="fp32")
    configure(s, WORKLOAD, solver_kind)
    for name, arr in seeded_config_state(WORKLOAD).items():
        s.upload(name, arr)
    return cfg, s


def time_gpu(solver_kind: str, steps: int, warmup: int):
    cfg, s = make_solver(solver_kind)
    dt = cfg["dt"]
    t_axis = [dt * i for i in range(warmup + steps)]
    if warmup:
        s.run(dt, t_axis[:warmup])
    t0 = time.perf_counter()
    frames = s.run(dt, t_axis[warmup:])  # returns after the stream has drained
    elapsed = time.perf_counter() - t0
    return cfg, s, elapsed, frames, s.run_iters()


def kernel_roofline(solver_kind: str, nvox: int, steps: int = 2):
    """Average device time of the CG kernels over `steps` steps, from HIP
    events recorded around each launch on the solver's stream.

    Algorithmic bytes (FP32 array passes of N cells): cg_dir reads r, d_old
    and writes d (3 passes); cg_update reads p, d, r and writes p, r (5
    passes); the one-launch iteration (k_cg_cgs4) counts the 8 passes of a
    whole CG iteration.  It physically moves 6.25 on average: it carries
    A r, rebuilds A d from the previous residual and touches d and p on
    every eighth launch only, which the measured traffic shows."""
    cfg, s = make_solver(solver_kind)
    s.set_kernel_timing(True)
    s.run(cfg["dt"], [cfg["dt"] * i for i in range(steps)])
    iters = s.run_iters()
    kern = {}
    for which, name, passes in ((0, "cg_dir", 3), (1, "cg_update", 5), (3, "k_cg_cgs4", 8)):
        avg, n = s.kernel_timing(which)
        if n == 0:
            continue
        bytes_ = passes * 4.0 * nvox
        kern[name] = {"avg_ms": avg, "launches": n, "bytes": bytes_, "total_ms": avg * n,
                      "gbps": bytes_ / (avg * 1e-3) / 1e9 if avg > 0 else 0.0}
    s.set_kernel_timing(False)
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    return dom, kern, iters


def load_traffic(kernel: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(seconds_budget: float = 20.0):
    """The f64 oracle (C restatement of src/fluid.rs) on this host: the
    OpenMP build on every core, and the single-thread build, on the first
    steps of the same workload."""
    from aether_fluid.scenarios import SMOKE_CONFIGS, seeded_config_state
    from oracle.oracle import OracleFluid
    cfg = SMOKE_CONFIGS[WORKLOAD]
    state = seeded_config_state(WORKLOAD)
    nvox = int(np.prod(cfg["dims"]))
    out = {}
    for omp in (True, False):
        o = OracleFluid(*cfg["dims"], h=cfg["h"], omp=omp)
        o.vorticity_eps, o.buoyancy_beta, o.gravity_alpha = cfg["eps"], cfg["beta"], cfg["alpha"]
        o.solver_kind, o.solver_iters, o.cg_rel_tol = 0, cfg["cg_max"], cfg["cg_tol"]
        o.set_source(1, cfg["source"]["center"], cfg["source"]["radius"], 1.0, 1.0)
        for name, arr in state.items():
            o.field(name)[...] = arr.reshape(o.nz, o.ny, o.nx)
        done, iters, t0 = 0, [], time.perf_counter()
        while True:
            o.step(cfg["dt"], cfg["dt"] * done)
            iters.append(o.stats()["iters"])
            done += 1
            el = time.perf_counter() - t0
            if el > seconds_budget / 2 or done >= 5:
                break
        out["omp" if omp else "single"] = {
            "value": nvox * done / el, "steps": done, "seconds": el,
            "threads": OracleFluid.num_threads(omp), "cg_iters": iters}
    return out


def _mixed_solver(prec: str, cg_tol: float = 0.0, cg_iters: int = 4000):
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, seeded_config_state
    cfg = SMOKE_CONFIGS["mixed_256"]
    s = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision=prec)
    for n, a in seeded_config_state("mixed_256").items():
        s.upload(n, a)
    s.gravity_alpha, s.buoyancy_beta, s.vorticity_eps = cfg["alpha"], cfg["beta"], cfg["eps"]
    s.set_sphere_source(cfg["source"]["center"], cfg["source"]["radius"])
    s.solver_kind, s.solver_iters, s.cg_rel_tol = 0, cg_iters, cg_tol
    if prec == "fp32_ir64":
        s.solver_iters = 1000
        s.ir_inner_rel_tol, s.ir_div_tol, s.ir_max_outer = 1e-3, 1e-7, 30
    return cfg, s


def _timed_steps(s, dt, t0_index, steps):
    """Runs `steps` steps one call each; returns (seconds, per-step max|div u|
    of the projection's FP64 residual or of the stored velocity, per-step CG
    iterations)."""
    el, divs, iters = 0.0, [], []
    for q in range(steps):
        t0 = time.perf_counter()
        s.run(dt, [dt * (t0_index + q)])
        el += time.perf_counter() - t0
        st = s.last_solve_stats()
        divs.append(st["div_max"] if st["outer"] > 0 else s.divergence_max())
        iters.append(s.run_iters()[0])
    return el, divs, iters


def bench_mixed(steps: int, warmup: int) -> dict:
    """256^3, dt 0.005: FP32 advection/stencils with the FP64-refined FP32
    CG projection (inner tol 1e-3, refined to max|div u| <= 1e-7 in FP64),
    against the all-FP64 projection stopped at the same divergence: FP64 CG
    to the loosest relative tolerance whose first step reaches
    max|div u| <= 1e-7 (found by trying 1e-8, 3e-9, 1e-9, ... on that step,
    outside the timed region), the same tolerance for every timed step."""
    cfg, _ = None, None
    nvox = 256 ** 3
    out = {"workload": "mixed-precision projection 256^3 against FP64 at the same max|div u|",
           "steps": steps, "warmup": warmup, "div_target": 1e-7}
    cfg, s = _mixed_solver("fp32_ir64")
    dt = cfg["dt"]
    if warmup:
        s.run(dt, [dt * q for q in range(warmup)])
    el, divs, its = _timed_steps(s, dt, warmup, steps)
    mixed_state = [s.download(n) for n in ("u", "v", "w", "density")]
    out["fp32_ir64"] = {"voxel_steps_per_s": nvox * steps / el, "ms_per_step": el * 1e3 / steps,
                        "solve_div_max": divs, "velocity_div_max": s.divergence_max(),
                        "inner_cg_iters": its}
    s.close()
    tau = None
    for t in (1e-8, 3e-9, 1e-9, 3e-10, 1e-10, 3e-11, 1e-11):
        _, c = _mixed_solver("fp64", cg_tol=t)
        c.run(dt, [0.0])
        ok = c.divergence_max() <= 1e-7
        c.close()
        if ok:
            tau = t
            break
    out["fp64_rel_tol"] = tau
    _, s = _mixed_solver("fp64", cg_tol=tau or 1e-11)
    if warmup:
        s.run(dt, [dt * q for q in range(warmup)])
    el64, divs64, its64 = _timed_steps(s, dt, warmup, steps)
    out["fp64"] = {"voxel_steps_per_s": nvox * steps / el64, "ms_per_step": el64 * 1e3 / steps,
                   "solve_div_max": divs64, "cg_iters": its64}
    fp64_state = [s.download(n) for n in ("u", "v", "w", "density")]
    s.close()
    out["rel_l2_mixed_vs_fp64"] = {
        n: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
        for n, x, y in zip(("u", "v", "w", "density"), mixed_state, fp64_state)}
    out["speedup_mixed_over_fp64"] = el64 / el
    return out


def precision_vs_fp64(steps: int, warmup: int, fp32_solver) -> dict:
    """The timed FP32 run against the FP64 instantiation (the reference's
    arithmetic; pinned to the f64 oracle bit for bit or to 1e-12) over the
    same steps from the same state: per-field relative L2 and max|div u|
    after the last projection of each.  Outside the timed region."""
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, configure, seeded_config_state
    cfg = SMOKE_CONFIGS[WORKLOAD]
    s64 = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision="fp64")
    configure(s64, WORKLOAD, "cg")
    for name, arr in seeded_config_state(WORKLOAD).items():
        s64.upload(name, arr)
    dt = cfg["dt"]
    s64.run(dt, [dt * i for i in range(warmup + steps)])
    out = {"reference": "FP64 instantiation of the same kernels, same steps",
           "steps": warmup + steps, "rel_l2_fp32_vs_fp64": {}}
    for n in ("u", "v", "w", "density", "temperature"):
        a, b = fp32_solver.download(n), s64.download(n)
        out["rel_l2_fp32_vs_fp64"][n] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out["div_max_fp32"] = fp32_solver.divergence_max()
    out["div_max_fp64"] = s64.divergence_max()
    out["cg_iters_fp64_last"] = s64.run_iters()[-1]
    s64.close()
    # one step from the seeded state in both precisions (the FP32 solver
    # starts from the FP32 rounding of it): the per-step difference, before
    # the flow's sensitivity amplifies it
    one = {}
    for prec in ("fp32", "fp64"):
        s1 = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision=prec)
        configure(s1, WORKLOAD, "cg")
        for name, arr in seeded_config_state(WORKLOAD).items():
            s1.upload(name, arr)
        s1.run(dt, [0.0])
        one[prec] = ({n: s1.download(n) for n in ("u", "v", "w", "density", "temperature")},
                     s1.divergence_max())
        s1.close()
    out["one_step"] = {
        "rel_l2_fp32_vs_fp64": {n: float(np.linalg.norm(one["fp32"][0][n] - one["fp64"][0][n]) /
                                         max(np.linalg.norm(one["fp64"][0][n]), 1e-300))
                                for n in one["fp64"][0]},
        "div_max_fp32": one["fp32"][1], "div_max_fp64": one["fp64"][1]}
    return out


def cg256_roofline(iters: int = 20) -> dict:
    """The HBM-bound case: one CG iteration at 256^3 FP32 (64 MiB per array,
    past the 256 MB Infinity Cache), in each layout, timed by the CG
    autotuner with HIP events over `iters` whole iterations (so the
    iteration time includes every launch of it: the direction and update
    kernels and, past 1024 partials, the k_partials_total pre-sum).
    Algorithmic bytes: 8 FP32 passes of N per iteration."""
    import aether_fluid as af
    n = 256
    s = af.FluidSolver3D(n, n, n, h=1.0 / n, precision="fp32")
    tune = s.autotune_cg(iters)
    s.close()
    bytes_it = 8 * 4.0 * n ** 3
    per = {k: {"us_per_iter": v, "gbps": bytes_it / (v * 1e-6) / 1e9}
           for k, v in tune["us_per_iter"].items()}
    best = min(per, key=lambda k: per[k]["us_per_iter"])
    # HBM-side bytes of one iteration of the four-wide pair with 4 planes per
    # thread, from the FETCH_SIZE / WRITE_SIZE passes in profiles/traffic.json
    traffic = None
    if best == "x4_z4":
        try:
            with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
                traffic = json.load(f).get("cg256_iteration", {}).get("bytes_per_iteration")
        except (OSError, ValueError):
            traffic = None
    return {"bound": "hbm", "kernel": f"CG iteration 256^3 ({best}: k_cg_dir4 + k_cg_update4 "
            "+ k_partials_total)" if best.startswith("x4") else f"CG iteration 256^3 ({best})",
            "achieved": per[best]["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": per[best]["gbps"] / HBM_PEAK_GBPS, "traffic": traffic,
            "algorithmic_bytes_per_iter": bytes_it, "layouts": per}


FD_EPS = (("fd_1e-3", 1e-3), ("fd_1e-4", 1e-4), ("fd_1e-5", 1e-5), ("fd_1e-6", 1e-6),
          ("fd_1e-7", 1e-7))


def bench_rollout(n: int, steps: int, ckpt: int, ndirs: int = 0, reps: int = 2,
                  init_state=None, cg_tol: float = 1e-5, cg_iters: int = 200,
                  fd_eps=FD_EPS, adj_tol: float | None = None) -> dict:
    """Differentiable rollout at n^3: FP32 forward, FP32 adjoint with FP64
    reductions, checkpoint every `ckpt` steps; inflow-patch controls and
    buoyancy beta; loss = sum (rho_T - target)^2.  Times the forward-only
    rollout against forward + adjoint, and (ndirs > 0) checks the gradient
    against the FP64 adjoint and FP64 central differences along seeded
    directions.  cg_tol / cg_iters: forward and adjoint projection solves."""
    import aether_fluid as af
    from aether_fluid.scenarios import rollout_problem
    pb = rollout_problem(n)
    s = af.FluidSolver3D(n, n, n, precision="fp32")
    s.gravity_alpha, s.vorticity_eps = 0.05, 0.3
    s.solver_kind, s.solver_iters, s.cg_rel_tol = 0, cg_iters, cg_tol
    if init_state is not None:
        for k, a in init_state.items():
            s.upload(k, a)
    args = dict(nsteps=steps, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
                target=pb["target"], checkpoint_every=ckpt)
    adj = {} if adj_tol is None else {"adjoint_cg_rel_tol": adj_tol}
    s.rollout(controls=pb["controls"], beta=1.0, grad=True, **args)  # warm-up
    t_f, t_fa, t_fs = [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        s.rollout(controls=pb["controls"], beta=1.0, grad=False, **args)
        t_f.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        loss, gc, gb = s.rollout(controls=pb["controls"], beta=1.0, grad=True, **args, **adj)
        t_fa.append(time.perf_counter() - t0)
        if adj:  # the adjoint solved to the forward's tolerance, for comparison
            t0 = time.perf_counter()
            s.rollout(controls=pb["controls"], beta=1.0, grad=True, **args)
            t_fs.append(time.perf_counter() - t0)
    tf, tfa = min(t_f), min(t_fa)
    nvox = n ** 3
    out = {"workload": f"differentiable rollout {n}^3, {steps} steps, checkpoint every {ckpt}",
           "cg_rel_tol": cg_tol, "cg_max_iters": cg_iters,
           "adjoint_cg_rel_tol": cg_tol if adj_tol is None else adj_tol,
           "fwd_voxel_steps_per_s": nvox * steps / tf,
           "fwd_adjoint_voxel_steps_per_s": nvox * steps / tfa,
           "fwd_adjoint_over_fwd": tfa / tf, "loss": loss, "grad_beta": gb}
    if t_fs:
        out["fwd_adjoint_over_fwd_adjoint_at_forward_tol"] = min(t_fs) / tf
    if ndirs:
        # The loss is piecewise smooth (MacCormack clamp, floor, near-wall
        # switches): central differences converge to the derivative only when
        # eps is small enough not to straddle kinks.  Report the FP32 adjoint
        # against (a) the FP64 adjoint and (b) FP64 central differences at the
        # specified eps = 1e-3 and at 1e-4 ... 1e-7.
        s64 = af.FluidSolver3D(n, n, n, precision="fp64")
        s64.gravity_alpha, s64.vorticity_eps = 0.05, 0.3
        s64.solver_kind, s64.solver_iters, s64.cg_rel_tol = 0, 3000, 1e-12
        if init_state is not None:
            for k, a in init_state.items():
                s64.upload(k, a)
        _, gc64, gb64 = s64.rollout(controls=pb["controls"], beta=1.0, grad=True, **args)
        # whole-gradient error of the FP32 adjoint against the FP64 adjoint
        out["fp32_grad_rel_l2_vs_fp64"] = float(np.linalg.norm(gc - gc64) / np.linalg.norm(gc64))
        out["fp32_grad_beta_rel_vs_fp64"] = float(abs(gb - gb64) / abs(gb64))
        rng = np.random.default_rng(0xAE7E0003)
        errs = {"fp64_adjoint": [], **{k: [] for k, _ in fd_eps}}
        fd64 = {k: [] for k, _ in fd_eps}
        for _ in range(ndirs):
            dc, db = rng.uniform(-1, 1, len(pb["controls"])), rng.uniform(-1, 1)
            ad = float(gc @ dc + gb * db)
            ad64 = float(gc64 @ dc + gb64 * db)
            errs["fp64_adjoint"].append(abs(ad - ad64) / abs(ad64))
            for key, eps in fd_eps:
                lp, _, _ = s64.rollout(controls=pb["controls"] + eps * dc, beta=1.0 + eps * db,
                                       grad=False, **args)
                lm, _, _ = s64.rollout(controls=pb["controls"] - eps * dc, beta=1.0 - eps * db,
                                       grad=False, **args)
                fd = (lp - lm) / (2 * eps)
                errs[key].append(abs(ad - fd) / abs(fd))
                fd64[key].append(abs(ad64 - fd) / abs(fd))
        out["fd_directions"] = ndirs
        out["fp32_adjoint_rel_err"] = {k: {"max": max(v), "median": float(np.median(v))}
                                       for k, v in errs.items()}
        out["fp64_adjoint_rel_err"] = {k: {"max": max(v), "median": float(np.median(v))}
                                       for k, v in fd64.items()}
    return out


def _batched_graph(af, insts, form, z, dt, times, warmup, steps, nvox, rho_ref, it_ref):
    """The batch with every step captured into a HIP graph and replayed,
    checked bit for bit against the stream-launched run (rho_ref, it_ref)."""
    bg = af.FluidBatch(insts, solver_iters=200, cg_rel_tol=1e-5, graph=True, cg_form=form,
                       cg_z_cells=z)
    if warmup:
        bg.run(dt, times[:warmup])
    t0 = time.perf_counter()
    bg.run(dt, times[warmup:])
    el_g = time.perf_counter() - t0
    same = bool(np.array_equal(bg.download(len(insts) - 1, "density"), rho_ref) and
                np.array_equal(bg.iterations(), it_ref))
    bg.close()
    return {"voxel_steps_per_s": nvox * steps / el_g, "ms_per_step": el_g * 1e3 / steps,
            "bitwise_equal_to_stream_launches": same,
            "note": "includes capturing two step graphs per run"}


def bench_batched(steps: int, warmup: int, dt: float = 0.01, cg: str = "auto",
                  graph: bool = True) -> dict:
    """64 grids drawn from {32^3, 48^3, 64^3, 96^3, 128x96x128}: one launch
    per stage for all of them, against the same grids stepped one at a time
    with the single-grid solver.  FP32, CG rel tol 1e-5 (max 200), zero
    initial state, sphere sources in the bottom quarter.  cg: the batch's CG
    form, "auto" for the batch autotuner's choice, else x4, scalar or fused
    (the one-launch form by groups).  graph=False skips the HIP-graph replay."""
    import aether_fluid as af
    insts = af.batched_instances(64)
    nvox = sum(d["nx"] * d["ny"] * d["nz"] for d in insts)
    warmup = max(warmup, 1)  # the autotuner solves a warm-up step's right-hand sides
    times = [dt * i for i in range(warmup + steps)]
    forms = {"x4": af.CG_X4, "scalar": af.CG_SCALAR, "fused": af.CG_FUSED}
    # CG layout of the batch, chosen by timing whole solves of a warm-up
    # step's right-hand sides, on a batch of its own, so that the timed batch
    # runs one layout from its first step (a forced form skips the tuning)
    tune_b = None
    if cg == "auto":
        tb = af.FluidBatch(insts, solver_iters=200, cg_rel_tol=1e-5)
        tb.run(dt, times[:warmup])
        tune_b = tb.autotune_cg(0)
        tb.close()
        form_b, z_b = forms[tune_b["choice"]], tune_b["z_cells"]
    else:
        form_b, z_b = forms[cg], (4 if cg != "scalar" else 0)
    b = af.FluidBatch(insts, solver_iters=200, cg_rel_tol=1e-5, cg_form=form_b,
                      cg_z_cells=z_b)
    b.run(dt, times[:warmup])
    t0 = time.perf_counter()
    b.run(dt, times[warmup:])
    el_b = time.perf_counter() - t0
    it_b = b.iterations()
    rho_b = b.download(len(insts) - 1, "density")
    b.close()
    out_graph = None
    if graph:  # the same batch with every step captured into a HIP graph and replayed
        out_graph = _batched_graph(af, insts, form_b, z_b, dt, times, warmup, steps, nvox,
                                   rho_b, it_b)
    el_1 = 0.0
    it_1 = []
    single_choices = {}
    for d in insts:
        s = af.FluidSolver3D(d["nx"], d["ny"], d["nz"])
        s.vorticity_eps, s.buoyancy_beta, s.gravity_alpha = (d["vorticity_eps"], 1.0,
                                                             d["gravity_alpha"])
        s.set_sphere_source(d["source_center"], d["source_radius"])
        s.solver_iters, s.cg_rel_tol = 200, 1e-5
        ch = s.autotune_cg(20)["choice"]  # each single grid tuned the same way
        key = f'{ch["form"]}_z{ch["z_cells"]}'
        single_choices[key] = single_choices.get(key, 0) + 1
        if warmup:
            s.run(dt, times[:warmup])
        t0 = time.perf_counter()
        s.run(dt, times[warmup:])
        el_1 += time.perf_counter() - t0
        it_1.append(s.run_iters())
        s.close()
    return {"workload": "64 batched smoke grids, FP32, CG rel tol 1e-5", "instances": 64,
            "total_voxels": nvox, "steps": steps, "warmup": warmup,
            "batched": {"voxel_steps_per_s": nvox * steps / el_b, "ms_per_step": el_b * 1e3 / steps,
                        "cg_iters_mean": float(np.mean(it_b)), "autotune": tune_b,
                        "cg_form": {v: k for k, v in forms.items()}[form_b]},
            "batched_graph": out_graph,
            "one_at_a_time": {"voxel_steps_per_s": nvox * steps / el_1,
                              "ms_per_step": el_1 * 1e3 / steps,
                              "cg_iters_mean": float(np.mean(it_1)),
                              "autotune_choices": single_choices},
            "speedup": el_1 / el_b}


def _medium_rays(n: int, seed: int):
    """n seeded rays from a sphere of radius 3 towards uniform points of the
    box [-1, 1]^3 (directions normalised), wavelengths in 380-750 nm and
    Xoshiro256++ states (any nonzero words)."""
    rng = np.random.default_rng(seed)
    org = rng.normal(size=(n, 3))
    org *= 3.0 / np.linalg.norm(org, axis=1)[:, None]
    d = rng.uniform(-1.0, 1.0, (n, 3)) - org
    d /= np.linalg.norm(d, axis=1)[:, None]
    lam = rng.uniform(380.0, 750.0, (n, 4))
    states = rng.integers(1, 1 << 62, size=(n, 4), dtype=np.uint64)
    return org, d, lam, states


MEDIUM_BOUNDS = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
MEDIUM_SCALE, MEDIUM_G = 2.5, 0.55


def bench_fp64(steps: int, warmup: int) -> dict:
    """The 128^3 workload in FP64, the reference's own precision: ms per step
    with the reference's default CG budget (32 iterations, src/fluid.rs:83)
    and with the workload's 200; the CG layouts' time per iteration (the
    autotuner's HIP-event timing of 20 whole iterations) at 128^3 and 256^3;
    red-black SOR time per sweep (60 sweeps, omega 1.8) in the default
    layout and one cell per thread, with a bitwise comparison of the two."""
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, configure, seeded_config_state
    cfg = SMOKE_CONFIGS[WORKLOAD]
    dt = cfg["dt"]
    state = seeded_config_state(WORKLOAD)
    out = {"workload": "plain smoke 128^3 in FP64", "steps": steps, "warmup": warmup}

    def solver(n=None):
        dims = cfg["dims"] if n is None else (n, n, n)
        s = af.FluidSolver3D(*dims, h=cfg["h"] if n is None else 1.0 / n, precision="fp64")
        if n is None:
            configure(s, WORKLOAD, "cg")
            for name, arr in state.items():
                s.upload(name, arr)
        return s

    for iters in (32, 200):
        s = solver()
        s.solver_iters = iters
        t_axis = [dt * i for i in range(warmup + steps)]
        if warmup:
            s.run(dt, t_axis[:warmup])
        t0 = time.perf_counter()
        s.run(dt, t_axis[warmup:])
        el = time.perf_counter() - t0
        out[f"cg_{iters}"] = {"ms_per_step": el * 1e3 / steps,
                              "voxel_steps_per_s": float(np.prod(cfg["dims"])) * steps / el}
        s.close()
    out["cg_us_per_iter"] = {}
    for n in (128, 256):
        s = solver(n)
        out["cg_us_per_iter"][f"{n}^3"] = s.autotune_cg(20)["us_per_iter"]
        s.close()
    sweeps = {}
    pres = {}
    for name, form in (("default", af.CG_AUTO), ("one_cell_per_thread", af.CG_SCALAR)):
        s = solver()
        s.solver_kind, s.solver_iters, s.sor_omega, s.cg_form = 1, 60, 1.8, form
        s.project()  # warm-up
        for k, arr in state.items():
            s.upload(k, arr)
        t0 = time.perf_counter()
        s.project()
        sweeps[name] = (time.perf_counter() - t0) * 1e6 / 60
        pres[name] = s.download("pressure")
        s.close()
    out["rbsor_us_per_sweep"] = sweeps
    out["rbsor_bitwise_equal"] = bool(np.array_equal(pres["default"], pres["one_cell_per_thread"]))
    return out


def bench_medium(s, nrays: int = 1 << 20) -> dict:
    """FluidMedium (src/fluid.rs:675-910) on the headline solver's final
    128^3 state: the device snapshot (from_solver), point sampling, and the
    path tracer's queries over nrays seeded rays, each with its own
    Xoshiro256++ state.  The API takes host arrays, so every rate includes
    the host <-> device copies of its inputs and results."""
    import aether_fluid as af
    t0 = time.perf_counter()
    med = af.FluidMedium.from_solver(s, MEDIUM_BOUNDS, MEDIUM_SCALE, MEDIUM_G)
    t_snap = time.perf_counter() - t0
    org, d, lam, st0 = _medium_rays(nrays, 0xAE7E0006)
    rays = af.FluidMedium.make_rays(org, d, 0.0, 10.0)
    out = {"workload": f"FluidMedium of the 128^3 state after the timed steps, bounds [-1, 1]^3, "
                       f"sigma_t scale {MEDIUM_SCALE}, g {MEDIUM_G}; {nrays} rays from radius 3 "
                       f"towards the box", "rays": nrays,
           "note": "rates include host<->device copies (the API takes host arrays)",
           "from_solver_ms": t_snap * 1e3, "max_density": med.max_density}
    pts = np.random.default_rng(7).uniform(-1.1, 1.1, (nrays, 3))

    def rate(fn):
        fn()  # first call: module load and allocation
        t = time.perf_counter()
        r = fn()
        return nrays / (time.perf_counter() - t), r

    out["sample_density_and_temp_points_per_s"], _ = rate(lambda: med.sample_density_and_temp(pts))
    out["sample_interaction_rays_per_s"], it = rate(
        lambda: med.sample_interaction(rays, lam, st0.copy()))
    out["transmittance_rays_per_s"], tr = rate(lambda: med.transmittance(rays, st0.copy()))
    out["sample_phase_hg_per_s"], _ = rate(lambda: med.sample_phase_hg(-d, st0.copy()))
    out["interaction_fraction"] = float(np.mean(it["hit"] == 1))
    out["mean_transmittance"] = float(np.mean(tr))
    med.close()
    return out


def medium_cpu_baseline(s, nrays: int = 1 << 16) -> dict:
    """The f64 oracle's FluidMedium queries on the first nrays rays of the
    same set, on one host thread, with the device snapshot's fields."""
    from oracle.oracle import OracleFluid
    nx, ny, nz = s.nx, s.ny, s.nz
    o = OracleFluid(nx, ny, nz)
    for f in ("density", "temperature"):
        o.field(f)[...] = s.download(f).reshape(nz, ny, nx)
    org, d, lam, st0 = _medium_rays(1 << 20, 0xAE7E0006)
    org, d, lam, st0 = org[:nrays], d[:nrays], lam[:nrays], st0[:nrays].copy()
    z, far = np.zeros(nrays), np.full(nrays, 10.0)
    t = time.perf_counter()
    o.medium_interaction(*MEDIUM_BOUNDS, MEDIUM_SCALE, MEDIUM_G, org, d, z, far, lam, st0.copy())
    t_i = time.perf_counter() - t
    t = time.perf_counter()
    o.medium_transmittance(*MEDIUM_BOUNDS, MEDIUM_SCALE, MEDIUM_G, org, d, z, far, st0.copy())
    t_t = time.perf_counter() - t
    return {"sample_interaction_rays_per_s": nrays / t_i, "transmittance_rays_per_s": nrays / t_t,
            "cores": 1, "kind": "port", "sample": f"the first {nrays} rays, f64 C oracle"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's own step count: 100 for "
                         "plain_128, 20 for mixed_256, 50 for batched_64)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the fwd+adjoint rollout")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the precision, 256^3 roofline, mixed-precision, 64^3 rollout "
                         "and batched sections of the default line")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--workload", default="plain_128",
                    choices=["plain_128", "mixed_256", "batched_64", "rollout_64", "medium_128",
                             "fp64_128"],
                    help="plain_128: the headline line; mixed_256: mixed-precision "
                         "projection at 256^3 against the all-FP64 projection; medium_128: "
                         "FluidMedium sampling of the 128^3 state after --steps steps; "
                         "fp64_128: the 128^3 workload in FP64 (step times, CG and RB-SOR "
                         "layouts)")
    ap.add_argument("--batch-cg", default="auto", choices=["auto", "x4", "scalar", "fused"],
                    help="batched_64: the batch's CG form (auto: the batch autotuner's choice)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched_64: skip the HIP-graph replay (stream launches only)")
    ap.add_argument("--probe", action="store_true",
                    help="counter-collection mode: 2 steps of the workload with each CG "
                         "layout and with RB-SOR, and a 256 MiB stream copy; no timing output")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"plain_128": 100, "mixed_256": 20, "batched_64": 50}.get(args.workload, 20)
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; the smoke step is not "
              "sharded across GPUs (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import aether_fluid as af
    if args.workload == "rollout_64":
        print(json.dumps(bench_rollout(64, 40, 8, ndirs=16)), flush=True)
        return
    if args.workload == "batched_64":
        print(json.dumps(bench_batched(args.steps, args.warmup, cg=args.batch_cg,
                                       graph=not args.no_graph)), flush=True)
        return
    if args.workload == "mixed_256":
        print(json.dumps(bench_mixed(args.steps, args.warmup)), flush=True)
        return
    if args.workload == "medium_128":
        _, s, _, _, _ = time_gpu("cg", args.steps, args.warmup)
        print(json.dumps(bench_medium(s)), flush=True)
        return
    if args.workload == "fp64_128":
        print(json.dumps(bench_fp64(args.steps, args.warmup)), flush=True)
        return
    if args.probe:  # short run for counter collection: both CG layouts, RB-SOR, the copy
        _, s = make_solver("cg")
        s.run(SMOKE_DT, [0.0, SMOKE_DT])
        _, s2 = make_solver("cg")
        s2.cg_form = af.CG_X4
        s2.run(SMOKE_DT, [0.0, SMOKE_DT])
        _, s3 = make_solver("rbgs")
        s3.run(SMOKE_DT, [0.0, SMOKE_DT])
        af.stream_copy_ms(COPY_BYTES, 5)
        print("probe done", flush=True)
        return

    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    note("headline 128^3 FP32 CG")
    cfg, s, elapsed, frames, run_iters = time_gpu("cg", args.steps, args.warmup)
    nvox = int(np.prod(cfg["dims"]))
    stats = s.last_solve_stats()
    value = nvox * args.steps / elapsed
    ms = elapsed * 1e3 / args.steps
    precision = None if args.no_extras else precision_vs_fp64(args.steps, args.warmup, s)

    # RBGS variant of the same workload (omega 1.8, 60 sweeps)
    _, _, el_rb, _, _ = time_gpu("rbgs", args.steps, args.warmup)

    dom, kern, timed_iters = kernel_roofline("cg", nvox)
    kd = kern[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": kd["gbps"], "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": kd["gbps"] / HBM_PEAK_GBPS,
                "traffic": load_traffic(dom), "avg_launch_ms": kd["avg_ms"],
                "algorithmic_bytes_per_launch": kd["bytes"], "launches_timed": kd["launches"],
                "cg_iters_timed": timed_iters,
                "other_kernels": {k: {"avg_ms": v["avg_ms"], "gbps": v["gbps"]}
                                  for k, v in kern.items() if k != dom}}

    copy_ms = af.stream_copy_ms(COPY_BYTES, 20)
    roofline["copy_gbps"] = 2.0 * COPY_BYTES / (copy_ms * 1e-3) / 1e9

    line = {
        "metric": METRIC, "value": value, "unit": "voxel-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded Xoshiro256++ initial state)",
        "config": {"workload": "plain smoke 128^3 FP32, CG rel tol 1e-5 (max 200)",
                   "grid": list(cfg["dims"]), "h": cfg["h"], "dt": cfg["dt"],
                   "solver": "cg", "parallelism": "single GPU"},
        "cg_iters_last_step": stats["iters"],
        "rbgs": {"value": nvox * args.steps / el_rb, "ms_per_step": el_rb * 1e3 / args.steps,
                 "sweeps": cfg["rbgs_sweeps"], "omega": cfg["sor_omega"],
                 "step_algorithmic_gbps": algorithmic_bytes_per_rbgs_step(
                     nvox, cfg["rbgs_sweeps"]) * args.steps / el_rb / 1e9},
        "roofline": roofline,
        "precision": precision,
        "fwd_adjoint": None,  # filled below
    }
    if not args.no_extras:
        note("256^3 CG iteration roofline")
        line["roofline_hbm_256"] = cg256_roofline()
        note("mixed precision 256^3 vs FP64")
        line["mixed_256"] = bench_mixed(3, 1)
        note("rollout 64^3 x 40")
        # forward projections to 1e-6 (the tight tolerance of the workload);
        # the adjoint's to 1e-5, which leaves the gradient unchanged (its
        # accuracy is set by the forward tolerance; checked against central
        # differences below)
        line["rollout_64"] = bench_rollout(64, 40, 8, ndirs=4, cg_tol=1e-6, cg_iters=1000,
                                           fd_eps=(("fd_1e-3", 1e-3), ("fd_1e-7", 1e-7)),
                                           adj_tol=1e-5, reps=3)
        note("batched 64 grids")
        line["batched_64"] = bench_batched(3, 1)
        note("FluidMedium sampling")
        line["medium_128"] = bench_medium(s)
        if not args.no_cpu:
            cpu_m = medium_cpu_baseline(s)
            line["medium_128"]["cpu_baseline"] = cpu_m
            line["medium_128"]["gpu_over_cpu_interaction"] = (
                line["medium_128"]["sample_interaction_rays_per_s"] /
                cpu_m["sample_interaction_rays_per_s"])
    if not args.no_adjoint:
        from aether_fluid.scenarios import seeded_config_state
        ro = bench_rollout(128, 8, 8, ndirs=0, reps=2,
                           init_state=seeded_config_state(WORKLOAD))
        line["fwd_adjoint"] = {"value": ro["fwd_adjoint_voxel_steps_per_s"],
                               "unit": "voxel-steps/s",
                               "fwd_value": ro["fwd_voxel_steps_per_s"],
                               "fwd_adjoint_over_fwd": ro["fwd_adjoint_over_fwd"],
                               "workload": "128^3 rollout, 8 steps, checkpoint every 8, "
                                           "inflow-patch controls + beta, FP32 adjoint"}
    if run_iters:
        line["cg_iters_mean"] = float(np.mean(run_iters))
        line["step_algorithmic_gbps"] = (sum(algorithmic_bytes_per_step(nvox, k)
                                             for k in run_iters) / elapsed / 1e9)
    if not args.no_cpu:
        note("CPU baseline")
        cb = cpu_baseline(args.cpu_seconds)
        omp = cb["omp"]
        line["cpu_baseline"] = {
            "value": omp["value"], "unit": "voxel-steps/s", "cores": omp["threads"],
            "kind": "port",
            "sample": (f"f64 C oracle (OpenMP) steps 1-{omp['steps']} of the same 128^3 "
                       f"workload, CG iters {omp['cg_iters']}"),
            "single_thread": {"value": cb["single"]["value"], "cores": 1,
                              "steps": cb["single"]["steps"]},
        }
        line["gpu_over_cpu"] = value / omp["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
