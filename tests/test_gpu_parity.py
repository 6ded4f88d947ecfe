"""GPU parity: the HIP library (through the C ABI) against the f64 oracle.

Bars:
* FP64 instantiation, stages without reductions (advection, wall BC,
  confinement/buoyancy, gradient, RBGS): bit-exact against the oracle.
* FP64 stages with reductions or device libm (CG, jets' exp, telemetry):
  relative error <= 1e-11 (fixed-order tree sums vs sequential sums).
* FP32 instantiation, same FP32 inputs on both sides: per-field relative L2
  <= 1e-5 after every stage of the 128^3 step, off the MacCormack floor and
  near-wall switch cells that the test finds from the velocity; CG
  iteration count within +-1 of the oracle's.  A whole unsynchronised step:
  1e-5 at 48^3, 1e-4 at 128^3 (one switch cell spread by the projection).
"""
import numpy as np
import pytest

from pairs import compare_fields, load_state, make_pair, random_state, rel_l2, set_knobs

pytestmark = pytest.mark.gpu

FP64_TIGHT = 1e-11


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_fp64_advection_bitexact(kind):
    _announce(f"fp64 advect kind={kind} 18x22x18")
    st = random_state(18, 22, 18, seed=10 + kind, vel=1.5)
    gpu, orc = make_pair(18, 22, 18, state=st)
    diss = 0.985 if kind == 4 else 0.996
    gpu.advect_maccormack_in_place(kind, 0.06, diss)
    orc.advect_maccormack_in_place(kind, 0.06, diss)
    name = ["u", "v", "w", "density", "temperature"][kind]
    np.testing.assert_array_equal(gpu.download(name), orc.field(name).ravel())


def test_fp64_wall_bc_and_confinement_bitexact():
    _announce("fp64 bc+confinement 16x20x16")
    st = random_state(16, 20, 16, seed=3)
    gpu, orc = make_pair(16, 20, 16, state=st)
    set_knobs(gpu, orc)
    gpu.enforce_wall_velocity_bc()
    orc.enforce_wall_velocity_bc()
    for n in ("u", "v", "w"):
        np.testing.assert_array_equal(gpu.download(n), orc.field(n).ravel())
    gpu.apply_buoyancy_and_vorticity_confinement(0.08)
    orc.apply_buoyancy_and_vorticity_confinement(0.08)
    for n in ("u", "v", "w"):
        np.testing.assert_array_equal(gpu.download(n), orc.field(n).ravel())


def test_fp64_jets_and_divergence():
    _announce("fp64 jets+divergence 16x20x16")
    st = random_state(16, 20, 16, seed=4)
    gpu, orc = make_pair(16, 20, 16, state=st)
    set_knobs(gpu, orc)
    gpu.inject_sources(0.5, 0.1)
    orc.inject_sources(0.5, 0.1)
    for n in ("u", "v", "w", "density", "temperature"):
        assert rel_l2(gpu.download(n), orc.field(n)) <= 1e-15
    g = gpu.compute_divergence_norm()
    o = orc.compute_divergence_norm()
    assert abs(g - o) <= FP64_TIGHT * o
    # the jets' exp() comes from the device libm (last-bit differences), so
    # the divergence of the forced field agrees to rounding, not bit for bit
    bo = orc.field("divergence").ravel()
    np.testing.assert_allclose(gpu.download("divergence"), bo, rtol=0,
                               atol=1e-14 * np.abs(bo).max())


def test_known_answer_readme_fp64():
    """README.md:54 known answer of the verify.rs:229-250 scenario."""
    _announce("known answer 16x20x16 fp64")
    gpu, orc = make_pair(16, 20, 16)
    set_knobs(gpu, orc, kind=0, iters=40)
    gpu.inject_cyberpunk_sources(0.5, 0.1)
    gpu.apply_buoyancy_and_vorticity_confinement(0.1)
    pre = gpu.compute_divergence_norm()
    gpu.project()
    post = gpu.compute_divergence_norm()
    assert f"{pre:.4e}" == "2.5751e+00"
    assert f"{post:.4e}" == "6.7289e-03"
    assert gpu.last_solve_stats()["iters"] == 40


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_projection_cg_and_rbgs_ratio(precision):
    """src/fluid.rs:918-956 on the GPU."""
    _announce(f"fluid.rs burst test {precision}")
    import aether_fluid as af
    for kind, iters in ((0, 45), (1, 90)):
        s = af.FluidSolver3D(18, 22, 18, precision=precision)
        s.solver_kind, s.solver_iters = kind, iters
        n = 18 * 22 * 18
        u, v, w = np.zeros(n), np.zeros(n), np.zeros(n)
        for k in range(6, 12):
            for j in range(5, 11):
                for i in range(6, 12):
                    idx = s.idx(i, j, k)
                    u[idx] = (i - 8.5) * 3.0
                    v[idx] = (j - 7.5) * 4.5
                    w[idx] = (k - 8.5) * 2.5
        s.u, s.v, s.w = u, v, w
        s.enforce_wall_velocity_bc()
        before = s.compute_divergence_norm()
        assert before > 1.0
        s.project()
        after = s.compute_divergence_norm()
        assert after < before * 0.02, (kind, before, after)


def test_fp64_cg_projection_parity():
    _announce("fp64 CG projection 18x22x18")
    st = random_state(18, 22, 18, seed=5)
    gpu, orc = make_pair(18, 22, 18, state=st)
    set_knobs(gpu, orc, kind=0, iters=60)
    gpu.project()
    orc.project()
    errs = compare_fields(gpu, orc, ("u", "v", "w", "pressure", "divergence"))
    assert max(errs.values()) <= FP64_TIGHT, errs
    gs, os_ = gpu.last_solve_stats(), orc.stats()
    assert gs["iters"] == os_["iters"]
    assert abs(gs["r2"] - os_["r2"]) <= 1e-8 * os_["r2"]


def test_fp64_cg_rel_tol_iteration_count():
    _announce("fp64 CG rel tol 24x24x24")
    st = random_state(24, 24, 24, seed=6)
    gpu, orc = make_pair(24, 24, 24, state=st)
    set_knobs(gpu, orc, kind=0, iters=200, rel_tol=1e-5)
    gpu.project()
    orc.project()
    gs, os_ = gpu.last_solve_stats(), orc.stats()
    assert os_["converged"] and gs["converged"]
    assert abs(gs["iters"] - os_["iters"]) <= 1
    errs = compare_fields(gpu, orc, ("u", "v", "w"))
    assert max(errs.values()) <= 1e-9, errs


@pytest.mark.parametrize("dims,zc", [((18, 22, 18), 4), ((64, 48, 40), 4), ((64, 48, 40), 2),
                                     ((36, 37, 41), 4), ((320, 24, 20), 4), ((6, 6, 6), 2),
                                     ((128, 20, 12), 4)])
def test_fp64_two_wide_cg_vs_oracle(dims, zc):
    """FP64 CG with two x-cells per thread (cg_form CG_X4 in FP64): the
    one-cell-per-thread form's per-cell arithmetic with the dot products
    summed in another order.  Rows of 3, 9, 18, 32, 64 and 160 lanes (two
    wavefront-wide segments and a partial one), 2 and 4 planes per thread.
    Odd and even iteration caps (an odd cap leaves p's last term to the tail
    kernel, a stop before it applies the term itself): iteration counts
    equal to the oracle's and fields within
    FP64_TIGHT; a tolerance stop: iterations within 1 and fields within 1e-9,
    as the one-cell-per-thread form."""
    import aether_fluid as af
    from pairs import load_state
    st = random_state(*dims, seed=7)
    gpu, orc = make_pair(*dims, state=st)
    gpu.cg_form = af.CG_X4
    gpu.cg_z_cells = zc
    for iters, tol in ((25, 0.0), (40, 0.0), (3000, 1e-5)):
        _announce(f"fp64 two-wide CG {dims} z{zc} iters {iters} tol {tol}")
        load_state(gpu, orc, st)
        set_knobs(gpu, orc, kind=0, iters=iters, rel_tol=tol)
        gpu.project()
        orc.project()
        gs, os_ = gpu.last_solve_stats(), orc.stats()
        names = ("u", "v", "w", "pressure", "divergence")
        errs = compare_fields(gpu, orc, names)
        if tol == 0.0:
            # (the 6^3 system reaches the absolute guard before either cap)
            assert gs["iters"] == os_["iters"], (gs, os_)
            assert max(errs.values()) <= FP64_TIGHT, errs
        else:
            assert os_["converged"] and gs["converged"]
            assert abs(gs["iters"] - os_["iters"]) <= 1, (gs, os_)
            assert max(errs.values()) <= 1e-9, errs


def test_fp64_two_wide_cg_past_4096_partials():
    """256 x 256 x 136 in FP64 (8.9 M cells): the two-wide tiling has 4352
    workgroups, so both launches read pre-summed partials.  Against the
    one-cell-per-thread FP64 form on the same state: equal iteration counts
    (an odd cap and a tolerance stop) and fields within 1e-12 relative."""
    import aether_fluid as af
    dims = (256, 256, 136)
    st = random_state(*dims, seed=9)
    out = {}
    for form in (af.CG_SCALAR, af.CG_X4):
        _announce(f"fp64 CG {dims} form {form}")
        s = af.FluidSolver3D(*dims, precision="fp64")
        s.solver_kind, s.cg_form = 0, form
        res = []
        for iters, tol in ((21, 0.0), (400, 1e-3)):
            for n, v in st.items():
                s.upload(n, v)
            s.solver_iters, s.cg_rel_tol = iters, tol
            s.project()
            res.append((s.last_solve_stats()["iters"],
                        {n: s.download(n) for n in ("u", "v", "w", "pressure")}))
        out[form] = res
        s.close()
    for (ia, fa), (ib, fb) in zip(out[af.CG_SCALAR], out[af.CG_X4]):
        assert ia == ib, (ia, ib)
        for n in fa:
            assert rel_l2(fb[n], fa[n]) <= 1e-12, n


def test_cg_zero_rhs_does_nothing():
    _announce("CG zero rhs")
    import aether_fluid as af
    s = af.FluidSolver3D(12, 12, 12, precision="fp64")
    s.solver_iters = 50
    s.project()
    st = s.last_solve_stats()
    assert st["iters"] == 0 and st["converged"]


def test_fp64_rbgs_bitexact():
    _announce("fp64 RBGS 18x22x18")
    st = random_state(18, 22, 18, seed=7)
    gpu, orc = make_pair(18, 22, 18, state=st)
    set_knobs(gpu, orc, kind=1, iters=30, omega=1.7)
    gpu.project()
    orc.project()
    for n in ("u", "v", "w", "pressure", "divergence"):
        np.testing.assert_array_equal(gpu.download(n), orc.field(n).ravel())


@pytest.mark.parametrize("precision,dims,sweeps", [
    ("fp32", (32, 24, 20), 7), ("fp32", (32, 24, 20), 8), ("fp32", (256, 10, 9), 5),
    ("fp64", (32, 24, 20), 7), ("fp64", (128, 12, 10), 8), ("fp64", (256, 10, 9), 5)])
def test_one_launch_sor_sweep_equals_colour_passes(precision, dims, sweeps):
    """RB-SOR where a sweep runs as one launch (colour-0 values formed where
    the colour-1 cells read them): FP32 rows of 8 and 64 float4 lanes, FP64
    rows of 16 and 64 double2 lanes (FP64 rows of 128 lanes run the two-wide
    colour passes), odd and even sweep counts.  The projection is bitwise
    equal to the one-cell-per-thread colour passes; FP32 is within its bar of
    the oracle, FP64 equal to it bit for bit."""
    _announce(f"one-launch SOR sweep {precision} {dims} x{sweeps}")
    import aether_fluid as af
    st = random_state(*dims, seed=41)
    a, orc = make_pair(*dims, precision=precision, state=st)
    set_knobs(a, orc, kind=1, iters=sweeps, omega=1.8)
    b = af.FluidSolver3D(*dims, precision=precision)
    for n, v in st.items():
        b.upload(n, v)
    b.solver_kind, b.solver_iters, b.sor_omega = 1, sweeps, 1.8
    b.cg_form = af.CG_SCALAR  # one cell per thread, two passes per sweep
    a.project()
    b.project()
    orc.project()
    for n in ("u", "v", "w", "pressure"):
        np.testing.assert_array_equal(a.download(n), b.download(n))
        if precision == "fp64":
            np.testing.assert_array_equal(a.download(n), orc.field(n).ravel(), err_msg=n)
    errs = compare_fields(a, orc, ("u", "v", "w", "pressure"))
    assert max(errs.values()) <= 1e-5, errs


def test_fp64_full_step_reference_defaults():
    """Two reference-default steps (jets, CG 32) as in src/fluid.rs:959-968."""
    _announce("fp64 full step 16x20x16 x2")
    gpu, orc = make_pair(16, 20, 16)
    set_knobs(gpu, orc)
    for t in (0.2, 0.4):
        tg = gpu.step(0.08, t).as_array()
        to = orc.step(0.08, t)
        np.testing.assert_allclose(tg, to, rtol=1e-10, atol=1e-300)
    errs = compare_fields(gpu, orc)
    assert max(errs.values()) <= 1e-10, errs
    d = gpu.download("density")
    assert np.all(np.isfinite(d)) and np.all(d >= 0)


@pytest.mark.parametrize("dims", [(6, 6, 6), (8, 6, 7), (12, 6, 9), (7, 9, 6)])
def test_smallest_grids_vs_oracle(dims):
    """Grids at the reference's lower limit (n >= 6, src/fluid.rs:54; a 4^3
    interior at 6^3) and rows of two or three float4 lanes: two FP64 steps
    with the reference defaults (jets, CG 32), then one FP32 projection per
    CG layout the grid allows, each against the oracle."""
    _announce(f"smallest grids {dims}")
    import aether_fluid as af
    st = random_state(*dims, seed=3)
    gpu, orc = make_pair(*dims, state=st)
    set_knobs(gpu, orc)
    for t in (0.2, 0.4):
        tg = gpu.step(0.08, t).as_array()
        to = orc.step(0.08, t)
        np.testing.assert_allclose(tg, to, rtol=1e-10, atol=1e-300)
    errs = compare_fields(gpu, orc)
    assert max(errs.values()) <= 1e-10, errs
    forms = [af.CG_SCALAR, af.CG_FUSED] + ([af.CG_X4] if dims[0] % 4 == 0 else [])
    for form in forms:
        g32, o = make_pair(*dims, precision="fp32", state=st)
        set_knobs(g32, o, iters=60, rel_tol=1e-4)
        g32.cg_form = form
        g32.project()
        o.project()
        sg, so = g32.last_solve_stats(), o.stats()
        assert abs(sg["iters"] - so["iters"]) <= 1, (form, sg, so)
        errs = compare_fields(g32, o, ("u", "v", "w", "pressure"))
        assert max(errs.values()) <= 1e-5, (form, errs)


@pytest.mark.parametrize("dims", [(15, 32, 9), (44, 9, 11), (36, 37, 41), (13, 41, 20),
                                  (28, 18, 18), (68, 12, 10)])
def test_assorted_grid_shapes_vs_oracle(dims):
    """Shapes drawn to miss the tile sizes (64 x 4 cells, 64-lane rows, 4 z
    planes per thread): two FP64 steps with a sphere source and the
    relative CG tolerance against the oracle, then one FP32 projection per
    CG layout the grid allows.

    FP32 bars: u, v, w within 1e-5 of the f64 oracle; the pressure within
    3e-5.  The pressure solves a singular Neumann system whose condition
    number reaches ~5e3 here (68 x 12 x 10: largest eigenvalue ~12, smallest
    nonzero ~(pi/66)^2), so FP32 rounding inside the 131 CG iterations moves
    it by ~1e-5 relative in every layout (measured 0.96-1.01e-5 across the
    three forms), while the velocities it corrects stay within ~1e-6."""
    _announce(f"assorted grid {dims}")
    import aether_fluid as af
    nx, ny, nz = dims
    st = random_state(*dims, seed=sum(dims))
    src = dict(center=(nx / 2.0, 2.5, nz / 2.0), radius=max(2.0, min(dims) / 4.0))
    gpu, orc = make_pair(*dims, state=st)
    set_knobs(gpu, orc, iters=200, rel_tol=1e-5, source=src)
    for t in (0.0, 0.05):
        tg = gpu.step(0.05, t).as_array()
        to = orc.step(0.05, t)
        np.testing.assert_allclose(tg, to, rtol=1e-10, atol=1e-300)
    assert gpu.last_solve_stats()["iters"] == orc.stats()["iters"]
    errs = compare_fields(gpu, orc)
    assert max(errs.values()) <= 1e-10, errs
    forms = [af.CG_SCALAR, af.CG_FUSED] + ([af.CG_X4] if nx % 4 == 0 else [])
    for form in forms:
        g32, o = make_pair(*dims, precision="fp32", state=st)
        set_knobs(g32, o, iters=150, rel_tol=1e-4)
        g32.cg_form = form
        g32.project()
        o.project()
        sg, so = g32.last_solve_stats(), o.stats()
        assert abs(sg["iters"] - so["iters"]) <= 1, (form, sg, so)
        errs = compare_fields(g32, o, ("u", "v", "w", "pressure"))
        assert max(errs["u"], errs["v"], errs["w"]) <= 1e-5, (form, errs)
        assert errs["pressure"] <= 3e-5, (form, errs)


def test_fp64_main_rs_stage_20_steps():
    """The src/main.rs:82-94 loop shape (eps 6.5, 28 iters, dt 0.08) at
    20x28x20 for 20 steps (tests/engine_integration.rs:18-22 grid)."""
    _announce("fp64 20 steps 20x28x20")
    gpu, orc = make_pair(20, 28, 20)
    set_knobs(gpu, orc, eps=6.5, iters=28)
    times = [15.0 * (i / 19.0) for i in range(20)]
    tg = np.array([f.as_array() for f in gpu.run(0.08, times)])
    to = np.array([orc.step(0.08, t) for t in times])
    # FP64 vs FP64: the CG dot products are summed in a different order, and
    # the nonlinear step (clamp switches, confinement) amplifies those 1e-16
    # differences by ~1e8 over 20 steps
    np.testing.assert_allclose(tg, to, rtol=1e-6, atol=1e-12)
    errs = compare_fields(gpu, orc)
    assert max(errs.values()) <= 1e-6, errs


def test_cli_selfcheck_and_fluid_stage(capsys):
    """The command line: the verify.rs:229-250 self-check reproduces the
    reference's printed digits (FP64) and passes its 3% criterion (FP64 and
    FP32); the main.rs:82-113 stage at its real grid (36x52x36) gives the
    oracle's telemetry (12 steps) and reports the medium's max density."""
    _announce("cli selfcheck + main.rs stage 36x52x36")
    from aether_fluid import cli
    from oracle.oracle import OracleFluid
    r = cli.selfcheck("fp64")
    assert r["passed"] and r["matches_reference_digits"], r
    assert cli.selfcheck("fp32")["passed"]
    assert cli.main(["selfcheck"]) == 0
    assert "[PASS] 3D Navier-Stokes CG Poisson Projection" in capsys.readouterr().out
    st = cli.fluid_stage(12, "fp64")
    orc = OracleFluid(36, 52, 36)
    orc.vorticity_eps, orc.solver_iters = 6.5, 28
    to = np.array([orc.step(0.08, 15.0 * (i / 11.0)) for i in range(12)])
    tg = np.array([f.as_array() for f in st["frames"]])
    np.testing.assert_allclose(tg, to, rtol=1e-6, atol=1e-12)
    assert st["peak_ke"] == tg[:, 1].max() and st["peak_vort"] == tg[:, 3].max()
    assert abs(st["max_density"] - orc.max_density()) <= 1e-6 * orc.max_density()
    assert "[1/4] 3D Navier-Stokes Solver (36x52x36, 12 steps)" in st["line"]


def _soundtrack_head():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "soundtrack_head.npz")
    with np.load(path) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_steps_reproduce_reference_soundtrack_head(precision):
    """renders/simulation_soundtrack.wav, first second, driven by the GPU's
    own telemetry: the first 9 steps of the main.rs:82-94 run (36x52x36,
    dt 0.08, eps 6.5, 28 CG iterations) through the audio.rs:397-639
    restatement with the two fitted normalisers. FP64: every stored sample
    within 1 LSB of 24-bit. FP32: telemetry stays within ~1e-6 relative
    through frame 4, so the first 0.5 s is within 16 LSB; from frame 4 on the
    flow amplifies any difference ~20x per step (FP64's ulp-level
    differences grow at the same rate, 1e-15 -> 1e-11 by frame 8), so the
    rest of the second only gets a loose 2^14 LSB bound."""
    _announce(f"{precision} main.rs run 36x52x36 x9 -> soundtrack head")
    import aether_fluid as af
    from oracle.oracle import soundtrack
    fx = _soundtrack_head()
    s = af.FluidSolver3D(36, 52, 36, precision=precision)
    s.vorticity_eps, s.solver_iters = 6.5, 28
    fr = np.array([f.as_array() for f in s.run(0.08, [15.0 * (i / 119) for i in range(9)])])
    fr = np.concatenate([fr, np.repeat(fr[-1:], 111, 0)])
    n, st = int(fx["nsamples"]), int(fx["stride"])
    pcm = soundtrack(fr, fx["norm"][0], fx["norm"][1], n)[::st].astype(np.int64)
    d = np.abs(pcm - fx["pcm_head"].astype(np.int64))
    print(f"[{precision}] max |diff| {d.max()} LSB, exact {np.mean(d == 0):.5f}", flush=True)
    if precision == "fp64":
        assert d.max() <= 1
    else:
        assert d[: 24000 // st].max() <= 16
        assert d.max() <= 1 << 14


def _plain_pair(name, dims, precision):
    from aether_fluid.scenarios import SMOKE_CONFIGS, seeded_config_state
    cfg = SMOKE_CONFIGS[name]
    state = seeded_config_state(name, dims=dims if dims != cfg["dims"] else None)
    gpu, orc = make_pair(*dims, h=cfg["h"] * cfg["dims"][0] / dims[0], precision=precision,
                         state=state)
    return cfg, gpu, orc


@pytest.mark.parametrize("solver", ["cg", "rbgs"])
def test_fp32_plain_smoke_one_step(solver):
    """FP32 device step vs the f64 oracle on the seeded 128^3 recipe at 48^3."""
    _announce(f"fp32 plain smoke 48^3 {solver}")
    cfg, gpu, orc = _plain_pair("plain_128", (48, 48, 48), "fp32")
    src = dict(center=(24.0, 4.0, 24.0), radius=4.5)
    if solver == "cg":
        set_knobs(gpu, orc, eps=0.3, beta=1.0, alpha=0.05, kind=0, iters=200, rel_tol=1e-5,
                  source=src)
    else:
        set_knobs(gpu, orc, eps=0.3, beta=1.0, alpha=0.05, kind=1, iters=60, omega=1.8,
                  source=src)
    gpu.step(0.01, 0.0)
    orc.step(0.01, 0.0)
    errs = compare_fields(gpu, orc, ("u", "v", "w", "density", "temperature"))
    print("rel L2", errs)
    assert max(errs.values()) <= 1e-5, errs
    if solver == "cg":
        assert abs(gpu.last_solve_stats()["iters"] - orc.stats()["iters"]) <= 1


def _plain_128_fp32_pair(solver):
    """The headline workload's FP32 solver and an f64 oracle (OpenMP build)
    with the same knobs, both holding the seeded state as the GPU stores it
    (FP32-rounded), so the two start from identical inputs."""
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, configure, seeded_config_state
    from oracle.oracle import OracleFluid
    cfg = SMOKE_CONFIGS["plain_128"]
    gpu = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision="fp32")
    configure(gpu, "plain_128", solver)
    orc = OracleFluid(*cfg["dims"], h=cfg["h"], omp=True)
    orc.vorticity_eps, orc.buoyancy_beta, orc.gravity_alpha = cfg["eps"], cfg["beta"], cfg["alpha"]
    if solver == "cg":
        orc.solver_kind, orc.solver_iters, orc.cg_rel_tol = 0, cfg["cg_max"], cfg["cg_tol"]
    else:
        orc.solver_kind, orc.solver_iters, orc.sor_omega = 1, cfg["rbgs_sweeps"], cfg["sor_omega"]
    orc.set_source(1, cfg["source"]["center"], cfg["source"]["radius"], 1.0, 1.0)
    for n, a in seeded_config_state("plain_128").items():
        gpu.upload(n, a)
    _sync_oracle(gpu, orc)
    return cfg, gpu, orc


_STATE = ("u", "v", "w", "density", "temperature", "pressure")


def _sync_oracle(gpu, orc):
    """Give the oracle the GPU's current state (FP32 values, exact in f64)."""
    for n in _STATE:
        orc.field(n)[...] = gpu.download(n).reshape(orc.nz, orc.ny, orc.nx)


@pytest.mark.parametrize("solver", ["cg", "rbgs"])
def test_fp32_plain_128_stages_vs_oracle_same_inputs(solver):
    """The 128^3 FP32 step, stage by stage, against the f64 oracle fed the
    GPU's own FP32 state before every stage (src/fluid.rs:598-611: advect
    u, v, w; sphere source; confinement and buoyancy; projection; advect
    density, temperature).  Bar: per-field relative L2 <= 1e-5 after every
    stage.  The MacCormack stages are compared off the cells whose trace
    lies within 1e-6 of a floor or near-wall switch (found from the shared
    velocity by pairs.advection_switch_cells): there a last-bit difference
    of the trace picks another corner cell, a discontinuity of the
    reference's scheme, not an error.  Every cell that differs by more
    than 1e-5 of the field's maximum must be one of those, and they must be
    few (primary switches < 1e-4 of the cells, with the cells that read
    them < 1e-3)."""
    from pairs import advection_switch_cells
    _announce(f"fp32 128^3 stages vs oracle on FP32 inputs ({solver})")
    cfg, gpu, orc = _plain_128_fp32_pair(solver)
    dt, dims, t = cfg["dt"], cfg["dims"], 0.0
    scale = dt / cfg["h"]
    ncell = int(np.prod(dims))
    report = {}

    def check(stage, names, mask=None):
        for n in names:
            a, b = gpu.download(n), orc.field(n).ravel()
            diff = np.abs(a - b)
            keep = np.ones(ncell, dtype=bool) if mask is None else ~mask.ravel()
            err = np.linalg.norm((a - b)[keep]) / max(np.linalg.norm(b[keep]), 1e-300)
            big = diff > 1e-5 * max(np.abs(b).max(), 1e-300)
            report[f"{stage}:{n}"] = (err, int(big.sum()), int((~keep).sum()))
            print(f"{stage:12s} {n:11s} rel L2 {err:.3e}  big {int(big.sum())}  "
                  f"excluded {int((~keep).sum())}", flush=True)
            assert err <= 1e-5, (stage, n, err)
            assert not (big & keep).any(), (stage, n, np.flatnonzero(big & keep)[:8])

    def switches():
        vel = [orc.field(n).ravel().copy() for n in ("u", "v", "w")]
        m, primary = advection_switch_cells(*vel, dims, scale)
        print(f"switch cells: {primary} primary, {int(m.sum())} excluded", flush=True)
        assert primary < 1e-4 * ncell and m.sum() < 1e-3 * ncell, (primary, m.sum())
        return m

    for kind, name, diss in ((0, "u", 0.996), (1, "v", 0.996), (2, "w", 0.996)):
        m = switches()
        gpu.advect_maccormack_in_place(kind, dt, diss)
        orc.advect_maccormack_in_place(kind, dt, diss)
        check(f"advect {name}", (name,), m)
        _sync_oracle(gpu, orc)
    gpu.inject_sources(t, dt)
    orc.inject_sources(t, dt)
    check("source", ("density", "temperature"))
    _sync_oracle(gpu, orc)
    gpu.apply_buoyancy_and_vorticity_confinement(dt)
    orc.apply_buoyancy_and_vorticity_confinement(dt)
    check("confine", ("u", "v", "w"))
    _sync_oracle(gpu, orc)
    gpu.project()
    orc.project()
    check("project", ("u", "v", "w"))
    if solver == "cg":
        assert abs(gpu.last_solve_stats()["iters"] - orc.stats()["iters"]) <= 1
    _sync_oracle(gpu, orc)
    for kind, name, diss in ((3, "density", 0.996), (4, "temperature", 0.985)):
        m = switches()
        gpu.advect_maccormack_in_place(kind, dt, diss)
        orc.advect_maccormack_in_place(kind, dt, diss)
        check(f"advect {name}", (name,), m)
        _sync_oracle(gpu, orc)


@pytest.mark.parametrize("solver", ["cg", "rbgs"])
def test_fp32_plain_128_one_step_vs_oracle(solver):
    """One whole FP32 step of the 128^3 workload against one f64 oracle step
    from the same FP32-rounded state, unsynchronised between stages.

    The same-input stage test above holds every stage to 1e-5 off the
    floor and near-wall switch cells.  Over a whole step a switch cell's
    difference (up to a few 1e-2 in that cell) passes through the
    projection, which spreads it over the grid, so the whole-step fields
    are bounded at 1e-4 per field; the telemetry at rtol 1e-4."""
    _announce(f"fp32 plain 128^3 one step vs oracle on FP32 inputs ({solver})")
    cfg, gpu, orc = _plain_128_fp32_pair(solver)
    tg = gpu.step(cfg["dt"], 0.0).as_array()
    to = orc.step(cfg["dt"], 0.0)
    errs = compare_fields(gpu, orc, ("u", "v", "w", "density", "temperature"))
    print("rel L2", errs, flush=True)
    assert max(errs.values()) <= 1e-4, errs
    np.testing.assert_allclose(tg, to, rtol=1e-4, atol=1e-12)
    if solver == "cg":
        assert abs(gpu.last_solve_stats()["iters"] - orc.stats()["iters"]) <= 1


@pytest.mark.parametrize("solver", ["rbgs", "cg"])
def test_fp64_plain_128_one_step_vs_oracle(solver):
    """The headline workload's step at its real size in the FP64
    instantiation, against one f64 oracle step (OpenMP build, whose RBGS
    equals the sequential sweep bit for bit) from the same seeded state.
    RB-SOR (60 sweeps): every field bit-identical.  CG (cap 200, reached):
    the same iteration count, fields within 1e-12 relative L2 (fixed-order
    device dot products against the oracle's)."""
    _announce(f"fp64 plain 128^3 one step vs oracle ({solver})")
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, configure, seeded_config_state
    from oracle.oracle import OracleFluid
    cfg = SMOKE_CONFIGS["plain_128"]
    state = seeded_config_state("plain_128")
    gpu = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision="fp64")
    configure(gpu, "plain_128", solver)
    orc = OracleFluid(*cfg["dims"], h=cfg["h"], omp=True)
    orc.vorticity_eps, orc.buoyancy_beta, orc.gravity_alpha = cfg["eps"], cfg["beta"], cfg["alpha"]
    if solver == "cg":
        orc.solver_kind, orc.solver_iters, orc.cg_rel_tol = 0, cfg["cg_max"], cfg["cg_tol"]
    else:
        orc.solver_kind, orc.solver_iters, orc.sor_omega = 1, cfg["rbgs_sweeps"], cfg["sor_omega"]
    orc.set_source(1, cfg["source"]["center"], cfg["source"]["radius"], 1.0, 1.0)
    load_state(gpu, orc, state)
    gpu.step(cfg["dt"], 0.0)
    orc.step(cfg["dt"], 0.0)
    names = ("u", "v", "w", "density", "temperature", "pressure")
    if solver == "rbgs":
        for n in names:
            np.testing.assert_array_equal(gpu.download(n), orc.field(n).ravel(), err_msg=n)
    else:
        assert gpu.last_solve_stats()["iters"] == orc.stats()["iters"]
        errs = compare_fields(gpu, orc, names)
        print("rel L2", errs, flush=True)
        assert max(errs.values()) <= 1e-12, errs


def test_fp32_cfg0_few_steps():
    """32x48x32, h = 1/32, zero start, sphere source: 5 FP32 steps."""
    _announce("fp32 32x48x32 5 steps")
    from aether_fluid.scenarios import SMOKE_CONFIGS
    cfg = SMOKE_CONFIGS["plain_32x48x32"]
    gpu, orc = make_pair(*cfg["dims"], h=cfg["h"], precision="fp32")
    set_knobs(gpu, orc, eps=0.3, beta=1.0, alpha=0.05, kind=0, iters=200, rel_tol=1e-5,
              source=cfg["source"])
    for s in range(5):
        gpu.step(cfg["dt"], 0.0)
        orc.step(cfg["dt"], 0.0)
        assert abs(gpu.last_solve_stats()["iters"] - orc.stats()["iters"]) <= 1
    errs = compare_fields(gpu, orc, ("u", "v", "w", "density", "temperature"))
    print("rel L2", errs)
    assert max(errs.values()) <= 1e-4, errs


@pytest.mark.parametrize("solver", ["rbgs", "cg"])
def test_fp64_cfg0_fifty_steps(solver):
    """The 32x48x32 workload as specified (h = 1/32, dt 0.02, zero start,
    sphere source r = 4 at (16, 4, 16), alpha 0.05, beta 1, eps 0.3), all 50
    steps in the FP64 instantiation against the f64 oracle.  RBGS (omega
    1.7, 40 sweeps) has no reductions: every field after 50 steps is
    bit-identical.  CG (relative tolerance 1e-5, max 200): iteration counts
    within +-1 every step, fields within 1e-12 relative L2 (the device's
    fixed-order dot products differ from the sequential ones in the last
    bits; observed ~1e-15 after 50 steps)."""
    _announce(f"fp64 32x48x32 50 steps {solver}")
    from aether_fluid.scenarios import SMOKE_CONFIGS
    cfg = SMOKE_CONFIGS["plain_32x48x32"]
    gpu, orc = make_pair(*cfg["dims"], h=cfg["h"], precision="fp64")
    if solver == "rbgs":
        set_knobs(gpu, orc, eps=0.3, beta=1.0, alpha=0.05, kind=1, iters=40, omega=1.7,
                  source=cfg["source"])
    else:
        set_knobs(gpu, orc, eps=0.3, beta=1.0, alpha=0.05, kind=0, iters=200, rel_tol=1e-5,
                  source=cfg["source"])
    worst = 0
    for step in range(50):
        gpu.step(cfg["dt"], cfg["dt"] * step)
        orc.step(cfg["dt"], cfg["dt"] * step)
        if solver == "cg":
            worst = max(worst, abs(gpu.last_solve_stats()["iters"] - orc.stats()["iters"]))
    assert worst <= 1, worst
    names = ("u", "v", "w", "density", "temperature", "pressure")
    if solver == "rbgs":
        for n in names:
            np.testing.assert_array_equal(gpu.download(n), orc.field(n).ravel(), err_msg=n)
    else:
        errs = compare_fields(gpu, orc, names)
        print("rel L2 after 50 steps", errs)
        assert max(errs.values()) <= 1e-12, errs


def test_determinism_bitwise():
    _announce("determinism 32^3")
    import aether_fluid as af
    outs = []
    for _ in range(2):
        s = af.FluidSolver3D(32, 32, 32, precision="fp32")
        st = random_state(32, 32, 32, seed=9)
        for n, a in st.items():
            s.upload(n, a)
        s.cg_rel_tol, s.solver_iters = 1e-5, 200
        s.run(0.02, [0.0, 0.1, 0.2])
        outs.append(np.concatenate([s.download(n) for n in ("u", "v", "w", "density",
                                                             "pressure")]))
    np.testing.assert_array_equal(outs[0], outs[1])


def test_sample_trilinear_and_max_density():
    _announce("sample + max density")
    st = random_state(12, 14, 10, seed=11)
    gpu, orc = make_pair(12, 14, 10, state=st)
    pts = np.random.default_rng(2).uniform(-1, 16, size=(64, 3))
    got = gpu.sample_trilinear("density", pts)
    exp = np.array([orc.sample_trilinear("density", *p) for p in pts])
    np.testing.assert_array_equal(got, exp)
    assert gpu.max_density() == max(0.0, st["density"].max())


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_nonfinite_flag(precision):
    """A NaN uploaded into one velocity cell is reported by the telemetry
    reduction's device flag (which a query clears); clean steps raise
    nothing."""
    _announce(f"non-finite flag {precision}")
    import aether_fluid as af
    dims = (20, 24, 16)
    s = af.FluidSolver3D(*dims, precision=precision)
    st = random_state(*dims, seed=4, vel=0.3)
    for n, a in st.items():
        s.upload(n, a)
    s.solver_iters = 40
    s.step(0.02, 0.0)
    assert not s.nonfinite()
    u = s.download("u")
    u[(8 * dims[1] + 10) * dims[0] + 7] = np.nan
    s.upload("u", u)
    s.compute_telemetry(0.0)
    assert s.nonfinite()
    assert not s.nonfinite()  # cleared by the query
    s.step(0.02, 0.02)  # the reference lets the NaN spread; the flag reports it
    assert s.nonfinite()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_sample_trilinear_caller_field(precision):
    """sample_trilinear on a caller's own f64 field (src/fluid.rs:94 takes
    any &[f64]): evaluated in FP64 whatever the solver's precision, so equal
    to the oracle bit for bit, with points inside, on and beyond the clamp
    box (src/fluid.rs:95-97)."""
    _announce(f"sample_trilinear on a caller field ({precision})")
    import aether_fluid as af
    from oracle.oracle import OracleFluid
    dims = (12, 14, 10)
    field = np.random.default_rng(5).uniform(-3, 3, int(np.prod(dims)))
    gpu = af.FluidSolver3D(*dims, precision=precision)
    orc = OracleFluid(*dims)
    orc.field("pressure")[...] = field.reshape(dims[2], dims[1], dims[0])
    pts = np.random.default_rng(6).uniform(-2, 15, size=(200, 3))
    pts[:4] = [[0.5, 0.5, 0.5], [10.499999, 12.499999, 8.499999], [3.0, 4.0, 5.0],
               [11.0, 13.0, 9.0]]
    got = gpu.sample_trilinear(field, pts)
    exp = np.array([orc.sample_trilinear("pressure", *p) for p in pts])
    np.testing.assert_array_equal(got, exp)
    with pytest.raises(af.AetherFluidError):
        gpu.sample_trilinear(field[:-1], pts)


def test_invalid_arguments_raise():
    import aether_fluid as af
    with pytest.raises(af.AetherFluidError):
        af.FluidSolver3D(5, 10, 10)
    s = af.FluidSolver3D(8, 8, 8)
    with pytest.raises(af.AetherFluidError):
        s.upload("u", np.zeros(7))
    with pytest.raises(af.AetherFluidError):
        s.advect_maccormack_in_place(7, 0.1, 1.0)
    # a count query may pass no array; an array length without one is an error
    import ctypes as C
    from aether_fluid import _lib as L
    cnt = C.c_int(-1)
    assert s._lib.aether_fluid_last_run_iterations(s._h, None, 0, C.byref(cnt)) == L.AETHER_OK
    assert cnt.value == 0
    assert s._lib.aether_fluid_last_run_iterations(s._h, None, 4, C.byref(cnt)) == L.AETHER_EINVAL


def _step_128(solver="cg"):
    """Two steps of the seeded 128^3 workload with the default CG layout (the
    one-launch four-wide form, whose queue-whole schedule the second solve
    takes after a capped first one)."""
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, configure, seeded_config_state
    cfg = SMOKE_CONFIGS["plain_128"]
    s = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision="fp32")
    configure(s, "plain_128", solver)
    for n, a in seeded_config_state("plain_128").items():
        s.upload(n, a)
    s.run(cfg["dt"], [0.0, cfg["dt"]])
    return s


def test_one_launch_cg_cycle_positions_vs_oracle():
    """The one-launch four-wide CG brings d and p up to date once every 8
    launches.  Iteration caps 16..23 end the solve at every position of the
    cycle (the tail kernel closes caps that are not multiples of it), and
    tolerances taken from the oracle's residual after 16..23 iterations stop
    it at every position (the stopping launch applies the pending terms).
    FP32 bar against the f64 oracle."""
    import aether_fluid as af
    from pairs import load_state
    cycle = 8
    dims = (64, 48, 40)
    state = random_state(*dims, seed=5)
    gpu, orc = make_pair(*dims, precision="fp32", state=state)
    gpu.cg_form = af.CG_FUSED
    for n in range(16, 24):
        for mode in ("cap", "tol"):
            iters, tol = n, 0.0
            if mode == "tol":
                load_state(gpu, orc, state)
                set_knobs(gpu, orc, iters=n, rel_tol=0.0)
                orc.project()
                so = orc.stats()
                iters, tol = 200, float(np.sqrt(so["r2"] / so["b2"])) * (1 + 1e-6)
            _announce(f"one-launch CG cycle {cycle} 64x48x40 {mode} {n}")
            load_state(gpu, orc, state)
            set_knobs(gpu, orc, iters=iters, rel_tol=tol)
            gpu.project()
            orc.project()
            so, sg = orc.stats(), gpu.last_solve_stats()
            if mode == "cap":
                assert sg["iters"] == so["iters"] == n, (sg, so)
            else:
                assert abs(so["iters"] - n) <= 1 and abs(sg["iters"] - so["iters"]) <= 1, (n, sg, so)
            errs = compare_fields(gpu, orc, ("u", "v", "w", "pressure"))
            assert max(errs.values()) <= 1e-5, (mode, n, errs)


@pytest.mark.parametrize("form", ["fused", "x4"])
def test_cg_queued_whole_after_capped_solve_equals_paced(form):
    """After a solve that ran to its cap the host queues the next solve's
    launches all at once instead of in paced chunks.  When that solve stops
    early (tolerance reached after ~11 iterations of a cap of 60), the
    launches queued after the stop must leave the result unchanged: fields
    and iteration count equal, bit for bit, those of a fresh handle whose
    solve runs paced."""
    _announce(f"CG queued whole after a capped solve ({form})")
    import aether_fluid as af
    dims = (64, 48, 40)
    state = random_state(*dims, seed=5)
    runs = []
    for history in (True, False):
        s = af.FluidSolver3D(*dims, precision="fp32")
        s.cg_form = af.CG_FUSED if form == "fused" else af.CG_X4
        s.solver_kind = 0
        if history:  # a capped solve first
            for n, a in state.items():
                s.upload(n, a)
            s.solver_iters, s.cg_rel_tol = 60, 0.0
            s.project()
            assert s.last_solve_stats()["iters"] == 60
        for n, a in state.items():
            s.upload(n, a)
        s.solver_iters, s.cg_rel_tol = 60, 3e-2
        s.project()
        runs.append((s.last_solve_stats()["iters"],
                     {n: s.download(n) for n in ("u", "v", "w", "pressure")}))
        s.close()
    (ia, fa), (ib, fb) = runs
    assert ia == ib and 0 < ia < 60, (ia, ib)
    for n in fa:
        np.testing.assert_array_equal(fa[n], fb[n], err_msg=n)


def test_one_launch_cg_segment_edges_vs_oracle():
    """The one-launch four-wide CG where rows span two 64-lane segments
    (320 cells: 64 + 16 float4 lanes), so x-neighbours across every wave
    boundary come from the edge loads, not from the adjacent lane: one
    projection against the f64 oracle (FP32 bar), and against the two-launch
    four-wide CG."""
    _announce("one-launch CG segment edges 320x24x20")
    import aether_fluid as af
    dims = (320, 24, 20)
    state = random_state(*dims, seed=11)
    gpu, orc = make_pair(*dims, precision="fp32", state=state)
    set_knobs(gpu, orc, iters=200, rel_tol=3e-4)  # the oracle converges in 112 iterations
    gpu.cg_form = af.CG_FUSED
    gpu.project()
    orc.project()
    so, sg = orc.stats(), gpu.last_solve_stats()
    assert 0 < so["iters"] < 200 and abs(sg["iters"] - so["iters"]) <= 1, (sg, so)
    errs = compare_fields(gpu, orc, ("u", "v", "w", "pressure"))
    assert max(errs.values()) <= 1e-5, errs
    x4 = af.FluidSolver3D(*dims, precision="fp32")
    x4.cg_form, x4.solver_iters, x4.cg_rel_tol = af.CG_X4, 200, 3e-4
    for name, arr in state.items():
        x4.upload(name, arr)
    x4.project()
    assert abs(x4.last_solve_stats()["iters"] - sg["iters"]) <= 1
    for n in ("u", "v", "w", "pressure"):
        assert rel_l2(gpu.download(n), x4.download(n)) <= 1e-5, n


def test_full_size_128_projection_properties():
    """Size-independent checks at 128^3: the recursive residual reached the
    reported value, the projected field's divergence equals -r/h^2 per cell
    (so max|div u| = max|r|/h^2), density stays in [0, 1], and a repeat run
    is bitwise identical."""
    _announce("128^3 properties (default one-launch CG)")
    s = _step_128()
    st = s.last_solve_stats()
    h = s.h
    div_max = s.divergence_max()
    assert st["iters"] > 0 and st["r2"] < st["r2_init"]
    # |div u|_max <= |r|_2 / h^2 (a cell value cannot exceed the 2-norm)
    assert div_max <= (st["r2"] ** 0.5) / (h * h) * 1.0001 + 1e-6
    d = s.download("density")
    assert d.min() >= 0.0 and d.max() <= 1.0 + 1e-6
    t = _step_128()
    assert s.run_iters() == t.run_iters()
    for n in ("u", "v", "w", "density", "temperature", "pressure"):
        np.testing.assert_array_equal(s.download(n), t.download(n), err_msg=n)


def _mixed_state(dims, seed=0xAE7E0002):
    from aether_fluid.scenarios import seeded_config_state
    return seeded_config_state("mixed_256", dims=dims, seed=seed)


def test_mixed_precision_projection_vs_oracle():
    """FP32 fields + FP64 iterative refinement (inner FP32 CG at 1e-3)
    reaches max|div u| <= 1e-7 in the FP64 residual and matches the f64
    oracle's tightly converged projection."""
    _announce("mixed precision projection 40^3")
    dims = (40, 40, 40)
    st = _mixed_state(dims)
    gpu, orc = make_pair(*dims, h=1.0 / 40, precision="fp32_ir64")
    for n, a in st.items():
        gpu.upload(n, a)
        orc.field(n)[...] = a.reshape(40, 40, 40)
    gpu.solver_kind, gpu.solver_iters = 0, 400
    gpu.ir_inner_rel_tol, gpu.ir_div_tol, gpu.ir_max_outer = 1e-3, 1e-7, 30
    orc.solver_kind, orc.solver_iters, orc.cg_rel_tol = 0, 5000, 1e-15
    gpu.project()
    orc.project()
    s = gpu.last_solve_stats()
    print("ir stats", s)
    assert s["converged"] and s["div_max"] <= 1e-7 and s["outer"] >= 2
    errs = compare_fields(gpu, orc, ("u", "v", "w"))
    assert max(errs.values()) <= 1e-6, errs
    p_g, p_o = gpu.download("pressure"), orc.field("pressure").ravel()
    inside = np.zeros((40, 40, 40), bool)
    inside[1:-1, 1:-1, 1:-1] = True
    inside = inside.ravel()
    dg, do = p_g[inside] - p_g[inside].mean(), p_o[inside] - p_o[inside].mean()
    # p is fixed by the solve to ~ max|r| relative to |b|: a few 1e-8
    assert rel_l2(dg, do) <= 1e-7


def test_mixed_precision_matches_fp64_projection_steps():
    """Three steps: the refined FP32 path against the all-FP64 GPU path run
    to the same divergence (FP64 CG at relative tolerance 1e-14)."""
    _announce("mixed vs fp64 steps 32^3")
    import aether_fluid as af
    dims = (32, 32, 32)
    st = _mixed_state(dims, seed=7)
    runs = {}
    for prec in ("fp32_ir64", "fp64"):
        s = af.FluidSolver3D(*dims, h=1.0 / 32, precision=prec)
        for n, a in st.items():
            s.upload(n, a)
        s.gravity_alpha, s.buoyancy_beta, s.vorticity_eps = 0.05, 1.0, 0.3
        s.set_sphere_source((16.0, 3.0, 16.0), 3.0)
        s.solver_kind, s.solver_iters = 0, 2000
        if prec == "fp64":
            s.cg_rel_tol = 1e-14
        else:
            s.ir_inner_rel_tol, s.ir_div_tol = 1e-3, 1e-7
        s.run(0.005, [0.0, 0.005, 0.01])
        runs[prec] = (s, s.last_solve_stats())
    a, b = runs["fp32_ir64"][0], runs["fp64"][0]
    assert runs["fp32_ir64"][1]["div_max"] <= 1e-7
    for n in ("u", "v", "w", "density"):
        assert rel_l2(a.download(n), b.download(n)) <= 1e-5, n


def test_mixed_precision_256_real_size_step():
    """The mixed-precision workload at its real size (256^3, dt 0.005, the
    seeded divergent state, sphere source r = 24): one step refines the FP32
    projection to max|div u| <= 1e-7 in the FP64 residual, and the velocity
    and density match the all-FP64 GPU step (FP64 CG to relative tolerance
    1e-14) within 1e-4 relative L2."""
    _announce("mixed vs fp64, one 256^3 step")
    import aether_fluid as af
    from aether_fluid.scenarios import SMOKE_CONFIGS, seeded_config_state
    cfg = SMOKE_CONFIGS["mixed_256"]
    assert tuple(cfg["dims"]) == (256, 256, 256)
    state = seeded_config_state("mixed_256")
    out = {}
    for prec in ("fp32_ir64", "fp64"):
        s = af.FluidSolver3D(*cfg["dims"], h=cfg["h"], precision=prec)
        for n, a in state.items():
            s.upload(n, a)
        s.gravity_alpha, s.buoyancy_beta, s.vorticity_eps = cfg["alpha"], cfg["beta"], cfg["eps"]
        s.set_sphere_source(cfg["source"]["center"], cfg["source"]["radius"])
        s.solver_kind, s.solver_iters = 0, 4000
        if prec == "fp64":
            s.cg_rel_tol = 1e-14
        else:
            s.solver_iters = 1000
            s.ir_inner_rel_tol, s.ir_div_tol, s.ir_max_outer = 1e-3, 1e-7, 30
        s.run(cfg["dt"], [0.0])
        out[prec] = (s.last_solve_stats(), {n: s.download(n) for n in ("u", "v", "w", "density")})
        s.close()
    st32, f32 = out["fp32_ir64"]
    print("ir stats", st32)
    assert st32["converged"] and st32["div_max"] <= 1e-7
    errs = {n: rel_l2(f32[n], out["fp64"][1][n]) for n in f32}
    print("mixed vs fp64", errs)
    assert max(errs.values()) <= 1e-4, errs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fluid_medium_vs_oracle(precision):
    """FluidMedium::from_solver and sample_density_and_temp (fluid.rs:703-788)
    on the device against the oracle: points inside, outside and on the
    walls of the bounds; the medium is a snapshot (a later solver step leaves
    it unchanged)."""
    _announce(f"fluid medium {precision} 14x18x12")
    import aether_fluid as af
    nx, ny, nz = 14, 18, 12
    st = random_state(nx, ny, nz, seed=77)
    gpu, orc = make_pair(nx, ny, nz, precision=precision, state=st)
    if precision == "fp32":  # the oracle sees the FP32-rounded fields the device holds
        for n in ("density", "temperature"):
            orc.field(n)[...] = gpu.download(n).reshape(nz, ny, nx)
    lo, hi = (-2.0, 0.5, 1.0), (2.0, 3.0, 2.5)
    rng = np.random.default_rng(8)
    pts = np.column_stack([rng.uniform(lo[a] - 0.2, hi[a] + 0.2, 400) for a in range(3)])
    pts[:8] = [lo, hi, (lo[0], 1.0, 2.0), (hi[0], 1.0, 2.0), (0.0, hi[1], 2.0), (0.0, 3.0, 1.75),
               (0.0, 2.9, 1.75), (0.0, 2.9875, 1.75)]
    med = af.FluidMedium.from_solver(gpu, (lo, hi), sigma_t_scale=3.5, phase_g=1.4)
    got = med.sample_density_and_temp(pts)
    exp = orc.medium_sample(lo, hi, pts)
    assert np.all((got == 0.0) == (exp == 0.0))
    np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-15)
    md = orc.max_density()
    assert med.max_density == md
    assert med.majorant_sigma_t == max(md * 3.5, 1e-4) and med.phase_g == 0.9
    gpu.gravity_alpha, gpu.solver_iters = 0.05, 5
    gpu.step(0.05, 0.0)
    np.testing.assert_array_equal(med.sample_density_and_temp(pts), got)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_autotune_keeps_state_and_results(precision):
    """The CG autotuner restores pressure and divergence, and the step after
    it (with whatever layout it chose) matches the untuned solver to the
    one-step FP32 bar (FP32) or to reduction-order differences (FP64)."""
    _announce(f"cg autotune {precision} 40x32x36")
    import aether_fluid as af
    nx, ny, nz = 40, 32, 36
    st = random_state(nx, ny, nz, seed=31)
    a = af.FluidSolver3D(nx, ny, nz, precision=precision)
    b = af.FluidSolver3D(nx, ny, nz, precision=precision)
    for s in (a, b):
        for n, v in st.items():
            s.upload(n, v)
        s.gravity_alpha, s.solver_iters, s.cg_rel_tol = 0.05, 120, 1e-6
        s.set_sphere_source((20.0, 4.0, 18.0), 3.0)
    p0, d0 = a.download("pressure"), a.download("divergence")
    r = a.autotune_cg(10)
    print("autotune", r)
    np.testing.assert_array_equal(a.download("pressure"), p0)
    np.testing.assert_array_equal(a.download("divergence"), d0)
    if precision == "fp64":  # nx % 2 == 0: the two-wide layout is a candidate too
        assert set(r["us_per_iter"]) == {"x4_z2", "x4_z4", "scalar_z4", "scalar_z8",
                                         "scalar_z16"}
    else:
        assert len(r["us_per_iter"]) == 6
    fa, fb = a.step(0.02, 0.0).as_array(), b.step(0.02, 0.0).as_array()
    assert abs(a.last_solve_stats()["iters"] - b.last_solve_stats()["iters"]) <= 1
    bar = 1e-5 if precision == "fp32" else 1e-9  # another layout sums its dots in another order
    for n in ("u", "v", "w", "density"):
        assert rel_l2(a.download(n), b.download(n)) <= bar, n
    assert rel_l2(a.download("pressure"), b.download("pressure")) <= 10 * bar
    np.testing.assert_allclose(fa, fb, rtol=bar)


def test_batch_autotune_keeps_state_and_results():
    """The batch autotuner restores every instance's pressure and divergence;
    the steps after it match an untuned batch (bit for bit when it keeps the
    four-wide layout with 4 planes per thread, to the FP32 bar otherwise)."""
    _announce("batch autotune")
    import aether_fluid as af
    insts = [dict(nx=32, ny=24, nz=20, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 10.0), source_radius=3.0),
             dict(nx=20, ny=28, nz=16, vorticity_eps=0.2, buoyancy_beta=1.0, gravity_alpha=0.08,
                  source_center=(10.0, 4.0, 8.0), source_radius=2.5)]
    bs = [af.FluidBatch(insts, solver_iters=80, cg_rel_tol=1e-5) for _ in range(2)]
    for b in bs:
        for q, d in enumerate(insts):
            st = random_state(d["nx"], d["ny"], d["nz"], seed=60 + q, vel=0.5)
            for n in ("u", "v", "w", "density", "temperature", "pressure"):
                b.upload(q, n, st[n])
    before = [bs[0].download(q, n) for q in range(2) for n in ("pressure", "divergence")]
    r = bs[0].autotune_cg(10)
    print("batch autotune", r)
    assert set(r["us_per_iter"]) == {"x4_z2", "x4_z4", "scalar_z8", "fused_z4"}
    assert (r["choice"], r["z_cells"]) in (("x4", 2), ("x4", 4), ("scalar", 8), ("fused", 4))
    after = [bs[0].download(q, n) for q in range(2) for n in ("pressure", "divergence")]
    for x, y in zip(before, after):
        np.testing.assert_array_equal(x, y)
    for b in bs:
        b.run(0.02, [0.0, 0.02])
    for q in range(2):
        for n in ("u", "v", "w", "density"):
            x, y = bs[0].download(q, n), bs[1].download(q, n)
            if r["choice"] == "x4" and r["z_cells"] == 4:
                np.testing.assert_array_equal(x, y)
            else:
                assert rel_l2(x, y) <= 1e-5, (q, n)


def test_batch_autotune_on_step_right_hand_sides():
    """autotune_cg(0) times whole solves of the last step's right-hand sides
    at the batch's tolerance and cap: refused before the first step; after
    it, pressure and divergence are restored, and the next step matches an
    untuned batch run alongside (bit for bit when it keeps the untuned
    four-wide layout with 4 planes per thread, to the FP32 bar otherwise)."""
    _announce("batch autotune on step right-hand sides")
    import aether_fluid as af
    insts = [dict(nx=32, ny=24, nz=20, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 10.0), source_radius=3.0),
             dict(nx=20, ny=28, nz=16, vorticity_eps=0.2, buoyancy_beta=1.0, gravity_alpha=0.08,
                  source_center=(10.0, 4.0, 8.0), source_radius=2.5)]
    bs = [af.FluidBatch(insts, solver_iters=80, cg_rel_tol=1e-5) for _ in range(2)]
    for b in bs:
        for q, d in enumerate(insts):
            st = random_state(d["nx"], d["ny"], d["nz"], seed=70 + q, vel=0.5)
            for n in ("u", "v", "w", "density", "temperature", "pressure"):
                b.upload(q, n, st[n])
    with pytest.raises(af.AetherFluidError):
        bs[0].autotune_cg(0)
    for b in bs:
        b.run(0.02, [0.0])
    before = [bs[0].download(q, n) for q in range(2) for n in ("pressure", "divergence")]
    r = bs[0].autotune_cg(0)
    print("batch autotune (step right-hand sides)", r)
    assert set(r["us_per_iter"]) == {"x4_z2", "x4_z4", "scalar_z8", "fused_z4"}
    after = [bs[0].download(q, n) for q in range(2) for n in ("pressure", "divergence")]
    for x, y in zip(before, after):
        np.testing.assert_array_equal(x, y)
    for b in bs:
        b.run(0.02, [0.02])
    for q in range(2):
        for n in ("u", "v", "w", "density"):
            x, y = bs[0].download(q, n), bs[1].download(q, n)
            if r["choice"] == "x4" and r["z_cells"] == 4:
                np.testing.assert_array_equal(x, y)
            else:
                assert rel_l2(x, y) <= 1e-5, (q, n)


@pytest.mark.parametrize("form", ["auto", "fused"])
def test_batched_graph_replay_equals_stream_launches(form):
    """Steps captured into HIP graphs and replayed give, bit for bit, what the
    same batch gives launched step by step: odd step counts (the density /
    temperature slot parity) and a second run continuing the first; with the
    untuned layout and with the one-launch form by groups (whose captured
    solve holds every group's launches, 60 of them, with a mid-cycle cap)."""
    _announce(f"batched graph replay vs stream launches ({form})")
    import aether_fluid as af
    cg_form = af.CG_AUTO if form == "auto" else af.CG_FUSED
    insts = [dict(nx=32, ny=24, nz=20, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 10.0), source_radius=3.0),
             dict(nx=20, ny=28, nz=16, vorticity_eps=0.2, buoyancy_beta=1.0, gravity_alpha=0.08,
                  source_center=(10.0, 4.0, 8.0), source_radius=2.5)]
    runs = []
    for graph in (False, True):
        b = af.FluidBatch(insts, solver_iters=60, cg_rel_tol=1e-5, graph=graph, cg_form=cg_form)
        for q, d in enumerate(insts):
            st = random_state(d["nx"], d["ny"], d["nz"], seed=90 + q, vel=0.5)
            for n in ("u", "v", "w", "density", "temperature"):
                b.upload(q, n, st[n])
        f1 = b.run(0.02, [0.0, 0.02, 0.04])
        it1 = b.iterations()
        f2 = b.run(0.02, [0.06, 0.08])
        it2 = b.iterations()
        fields = [b.download(q, n) for q in range(2)
                  for n in ("u", "v", "w", "density", "temperature", "pressure")]
        tel = np.array([[fr.as_array() for fr in row] for row in f1 + f2])
        runs.append((fields, tel, it1, it2))
    (fa, ta, ia1, ia2), (fb, tb, ib1, ib2) = runs
    np.testing.assert_array_equal(ia1, ib1)
    np.testing.assert_array_equal(ia2, ib2)
    np.testing.assert_array_equal(ta, tb)
    for x, y in zip(fa, fb):
        np.testing.assert_array_equal(x, y)


def _batched_vs_single(insts, single_form, case, cap=120, tol=1e-5, z_cells=0, batch_form=0):
    """A batch of grids of different sizes gives, bit for bit, the fields and
    CG iteration counts of single-grid solvers run one at a time with the
    same CG layout and tiling (z_cells; 8 z cells per thread in the
    initialisation tiling).  batch_form: the batch's cg_form (0: AUTO)."""
    _announce(case)
    import os
    import aether_fluid as af
    b = af.FluidBatch(insts, solver_iters=cap, cg_rel_tol=tol, cg_z_cells=z_cells,
                      cg_form=batch_form)
    states = []
    for q, d in enumerate(insts):
        st = random_state(d["nx"], d["ny"], d["nz"], seed=40 + q, vel=0.5)
        states.append(st)
        for n in ("u", "v", "w", "density", "temperature"):
            b.upload(q, n, st[n])
    times = [0.0, 0.02, 0.04]
    frames = b.run(0.02, times)
    iters = b.iterations()
    os.environ["AETHER_CG_ZU"] = "8"
    try:
        for q, d in enumerate(insts):
            s = af.FluidSolver3D(d["nx"], d["ny"], d["nz"])
            s.vorticity_eps, s.buoyancy_beta, s.gravity_alpha = (d["vorticity_eps"], 1.0,
                                                                 d["gravity_alpha"])
            s.set_sphere_source(d["source_center"], d["source_radius"])
            s.solver_iters, s.cg_rel_tol = cap, tol
            s.cg_form = single_form
            s.cg_z_cells = z_cells
            for n in ("u", "v", "w", "density", "temperature"):
                s.upload(n, states[q][n])
            single = s.run(0.02, times)
            assert list(iters[:, q]) == s.run_iters(), (q, iters[:, q], s.run_iters())
            for n in ("u", "v", "w", "density", "temperature", "pressure"):
                np.testing.assert_array_equal(b.download(q, n), s.download(n), err_msg=f"{q} {n}")
            for t in range(len(times)):
                np.testing.assert_allclose(frames[t][q].as_array(), single[t].as_array(),
                                           rtol=1e-12, atol=1e-300)
    finally:
        os.environ.pop("AETHER_CG_ZU", None)


def test_batched_equals_one_at_a_time_bitwise():
    """nx = 70 is not a multiple of 4, so the batch runs the one-cell-per-
    thread CG; compared with single grids pinned to that layout."""
    import aether_fluid as af
    insts = [dict(nx=32, ny=32, nz=32, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 16.0), source_radius=3.5),
             dict(nx=24, ny=40, nz=20, vorticity_eps=0.2, buoyancy_beta=1.0, gravity_alpha=0.08,
                  source_center=(12.0, 5.0, 10.0), source_radius=3.0),
             dict(nx=70, ny=20, nz=18, vorticity_eps=0.45, buoyancy_beta=1.0, gravity_alpha=0.03,
                  source_center=(30.0, 3.0, 9.0), source_radius=2.5),
             dict(nx=16, ny=16, nz=40, vorticity_eps=0.1, buoyancy_beta=1.0, gravity_alpha=0.1,
                  source_center=(8.0, 3.0, 20.0), source_radius=2.0)]
    _batched_vs_single(insts, af.CG_SCALAR, "batched vs one-at-a-time (scalar CG)")


def test_batched_four_wide_equals_one_at_a_time_bitwise():
    """Every nx % 4 == 0 (including row widths of 6 and 17 float4 lanes, whose
    workgroups leave threads idle), so the batch runs the four-wide CG."""
    import aether_fluid as af
    insts = [dict(nx=32, ny=32, nz=32, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 16.0), source_radius=3.5),
             dict(nx=24, ny=40, nz=20, vorticity_eps=0.2, buoyancy_beta=1.0, gravity_alpha=0.08,
                  source_center=(12.0, 5.0, 10.0), source_radius=3.0),
             dict(nx=68, ny=20, nz=18, vorticity_eps=0.45, buoyancy_beta=1.0, gravity_alpha=0.03,
                  source_center=(30.0, 3.0, 9.0), source_radius=2.5),
             dict(nx=16, ny=16, nz=40, vorticity_eps=0.1, buoyancy_beta=1.0, gravity_alpha=0.1,
                  source_center=(8.0, 3.0, 20.0), source_radius=2.0)]
    _batched_vs_single(insts, af.CG_X4, "batched vs one-at-a-time (four-wide CG)")
    # an odd cap with no tolerance ends every solve after an even iteration,
    # whose alpha d term the batch's tail kernel applies
    _batched_vs_single(insts, af.CG_X4, "batched vs one-at-a-time (four-wide CG, cap 7)",
                       cap=7, tol=0.0)


def test_batched_four_wide_two_planes_equals_one_at_a_time_bitwise():
    """The four-wide CG with 2 z planes per thread (cg_z_cells = 2, a tiling
    the batch autotuner also times): the batch equals single grids pinned to
    the same layout and tiling, also with an odd cap (the tail kernel); a
    tiling the layout does not take is rejected."""
    import aether_fluid as af
    insts = [dict(nx=32, ny=32, nz=32, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 16.0), source_radius=3.5),
             dict(nx=68, ny=20, nz=19, vorticity_eps=0.45, buoyancy_beta=1.0, gravity_alpha=0.03,
                  source_center=(30.0, 3.0, 9.0), source_radius=2.5),
             dict(nx=16, ny=16, nz=40, vorticity_eps=0.1, buoyancy_beta=1.0, gravity_alpha=0.1,
                  source_center=(8.0, 3.0, 20.0), source_radius=2.0)]
    _batched_vs_single(insts, af.CG_X4, "batched vs one-at-a-time (four-wide CG, 2 planes)",
                       z_cells=2)
    _batched_vs_single(insts, af.CG_X4, "batched vs one-at-a-time (four-wide CG, 2 planes, cap 7)",
                       cap=7, tol=0.0, z_cells=2)
    b = af.FluidBatch(insts, solver_iters=10, cg_z_cells=16)
    with pytest.raises(af.AetherFluidError, match="cg_z_cells"):
        b.run(0.02, [0.0])
    b.close()


@pytest.mark.parametrize("form", ["x4", "fused"])
def test_batched_64_workload_equals_one_at_a_time_bitwise(form):
    """The batched workload's own 64 seeded grids (sizes from 32^3, 48^3,
    64^3, 96^3 and 128x96x128; 31.3 M voxels; seeded sources and forcing),
    three steps with CG to relative tolerance 1e-5: the batch (four-wide CG,
    or the one-launch CG solved group by group) equals the grids run one at
    a time with the same layout, bit for bit."""
    import aether_fluid as af
    insts = af.batched_instances(64)
    assert sum(d["nx"] * d["ny"] * d["nz"] for d in insts) > 30_000_000
    f = af.CG_X4 if form == "x4" else af.CG_FUSED
    _batched_vs_single(insts, f, f"batched vs one-at-a-time, the 64-grid workload ({form})",
                       cap=200, tol=1e-5, batch_form=af.CG_AUTO if form == "x4" else f)


def test_batched_one_launch_equals_one_at_a_time_bitwise():
    """The one-launch CG for batches (cg_form CG_FUSED), solved group by
    group: instances of every tiling class (two 128-cell rows per wavefront,
    regular tilings of 32 and 64 cells per row, irregular rows of 6 and 17
    float4 lanes) equal single grids pinned to the one-launch layout, bit for
    bit, with a tolerance that stops instances of one group at different
    launches (the converged ones return early from the rest), and with a cap
    of 13 that ends every solve mid-cycle (the per-instance tail)."""
    import aether_fluid as af
    insts = [dict(nx=128, ny=16, nz=16, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(64.0, 4.0, 8.0), source_radius=3.0),
             dict(nx=32, ny=32, nz=32, vorticity_eps=0.3, buoyancy_beta=1.0, gravity_alpha=0.05,
                  source_center=(16.0, 4.0, 16.0), source_radius=3.5),
             dict(nx=24, ny=40, nz=20, vorticity_eps=0.2, buoyancy_beta=1.0, gravity_alpha=0.08,
                  source_center=(12.0, 5.0, 10.0), source_radius=3.0),
             dict(nx=68, ny=20, nz=18, vorticity_eps=0.45, buoyancy_beta=1.0, gravity_alpha=0.03,
                  source_center=(30.0, 3.0, 9.0), source_radius=2.5),
             dict(nx=64, ny=16, nz=24, vorticity_eps=0.1, buoyancy_beta=1.0, gravity_alpha=0.1,
                  source_center=(32.0, 3.0, 12.0), source_radius=2.0),
             dict(nx=32, ny=32, nz=16, vorticity_eps=0.25, buoyancy_beta=1.0, gravity_alpha=0.06,
                  source_center=(16.0, 4.0, 8.0), source_radius=2.5)]
    _batched_vs_single(insts, af.CG_FUSED, "batched one-launch CG vs one-at-a-time",
                       batch_form=af.CG_FUSED)
    _batched_vs_single(insts, af.CG_FUSED, "batched one-launch CG vs one-at-a-time (cap 13)",
                       cap=13, tol=0.0, batch_form=af.CG_FUSED)


def _rollout_setup(n, precision, tol=1e-12, iters=600):
    import aether_fluid as af
    from aether_fluid.scenarios import rollout_problem
    prob = rollout_problem(n)
    s = af.FluidSolver3D(n, n, n, precision=precision)
    s.gravity_alpha, s.vorticity_eps = 0.05, 0.3
    s.solver_kind, s.solver_iters, s.cg_rel_tol = 0, iters, tol
    return s, prob


def _directions(count, nctrl, seed=123):
    rng = np.random.default_rng(seed)
    return [(rng.uniform(-1, 1, nctrl), rng.uniform(-1, 1)) for _ in range(count)]


def test_adjoint_fp64_matches_central_differences():
    """FP64 forward + FP64 adjoint against FP64 central differences of the
    loss along random directions (controls and beta)."""
    _announce("adjoint fp64 vs FD 16^3")
    n, steps = 16, 6
    s, pb = _rollout_setup(n, "fp64")
    beta = 1.0
    args = dict(nsteps=steps, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
                target=pb["target"], checkpoint_every=2)
    loss, gc, gb = s.rollout(controls=pb["controls"], beta=beta, **args)
    assert loss > 0 and np.all(np.isfinite(gc))
    eps = 1e-6
    for dc, db in _directions(4, len(pb["controls"])):
        lp, _, _ = s.rollout(controls=pb["controls"] + eps * dc, beta=beta + eps * db,
                             grad=False, **args)
        lm, _, _ = s.rollout(controls=pb["controls"] - eps * dc, beta=beta - eps * db,
                             grad=False, **args)
        fd = (lp - lm) / (2 * eps)
        ad = float(gc @ dc + gb * db)
        print("fp64 dir", fd, ad)
        assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12, (ad, fd)


def test_adjoint_fp32_matches_fp64_fd_and_checkpoints():
    """FP32 adjoint (FP64 reductions) against FP64 central differences
    (eps = 1e-3), and the gradient is bit for bit independent of the
    checkpoint interval; the solver state is left unchanged."""
    _announce("adjoint fp32 vs fp64 FD 24^3")
    n, steps = 24, 8
    s32, pb = _rollout_setup(n, "fp32", tol=1e-7, iters=800)
    s64, _ = _rollout_setup(n, "fp64")
    u_before = s32.download("u")
    args = dict(nsteps=steps, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
                target=pb["target"])
    loss, gc, gb = s32.rollout(controls=pb["controls"], beta=1.0, checkpoint_every=4, **args)
    _, gc1, gb1 = s32.rollout(controls=pb["controls"], beta=1.0, checkpoint_every=1, **args)
    np.testing.assert_array_equal(gc1, gc)  # fixed-point scatters: no order dependence
    assert gb1 == gb
    np.testing.assert_array_equal(s32.download("u"), u_before)
    eps = 1e-3
    worst = 0.0
    for dc, db in _directions(4, len(pb["controls"]), seed=7):
        lp, _, _ = s64.rollout(controls=pb["controls"] + eps * dc, beta=1.0 + eps * db,
                               grad=False, checkpoint_every=4, **args)
        lm, _, _ = s64.rollout(controls=pb["controls"] - eps * dc, beta=1.0 - eps * db,
                               grad=False, checkpoint_every=4, **args)
        fd = (lp - lm) / (2 * eps)
        ad = float(gc @ dc + gb * db)
        worst = max(worst, abs(ad - fd) / abs(fd))
        print("fp32 dir", fd, ad)
    assert worst <= 1e-3, worst


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("n", [24, 96])
def test_adjoint_bitwise_reproducible(precision, n):
    """The adjoint's scatters accumulate in 64-bit fixed point, so the
    gradient does not depend on thread order: two identical rollouts, and
    checkpoints every step against every 4 steps, give the same bits.  96^3
    runs the LDS-windowed scatter kernels (at least 256 tiles), 24^3 the
    one-cell-per-thread ones."""
    _announce(f"adjoint bitwise reproducible {precision} {n}^3")
    s, pb = _rollout_setup(n, precision, tol=1e-6, iters=400)
    args = dict(nsteps=4, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
                target=pb["target"])
    runs = [s.rollout(controls=pb["controls"], beta=1.0, checkpoint_every=k, **args)
            for k in (4, 4, 1)]
    for loss, gc, gb in runs[1:]:
        assert loss == runs[0][0]
        np.testing.assert_array_equal(gc, runs[0][1])
        assert gb == runs[0][2]
    assert np.abs(runs[0][1]).max() > 0


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_rollout_vjp_reproduces_fused_loss_gradient(precision):
    """The reverse-mode node: rollout_forward returns the final state, and
    rollout_backward with the adjoint 2 (density_T - target) computed on the
    host gives rollout()'s gradient bit for bit.  At 96^3 (3456 reduction
    workgroups, more than the 1024 slots most partial buffers hold) the
    fused loss equals sum (density_T - target)^2 of the downloaded final
    state."""
    _announce(f"rollout vjp vs fused loss gradient {precision} 96^3")
    s, pb = _rollout_setup(96, precision, tol=1e-6, iters=400)
    kw = dict(nsteps=3, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
              checkpoint_every=2)
    u0 = s.download("u")
    loss, gc, gb = s.rollout(controls=pb["controls"], beta=1.0, target=pb["target"], **kw)
    final = s.rollout_forward(controls=pb["controls"], beta=1.0, **kw)
    d = final["density"] - pb["target"]
    host_loss = float(np.sum(d * d))
    print("loss", loss, host_loss, flush=True)
    assert abs(loss - host_loss) <= 1e-12 * abs(host_loss)
    g2c, g2b, g2i = s.rollout_backward({"density": 2.0 * d})
    np.testing.assert_array_equal(g2c, gc)
    assert g2b == gb
    np.testing.assert_array_equal(s.download("u"), u0)  # state left as it was
    # the inflow patch is overwritten by the controls: no initial adjoint there
    lo, hi = pb["patch_lo"], pb["patch_hi"]
    rho0 = g2i["density"].reshape(96, 96, 96)
    assert np.all(rho0[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] == 0.0)
    import aether_fluid as af
    with pytest.raises(af.AetherFluidError):
        s.rollout_backward({"density": 2.0 * d})  # the tape was consumed


def test_rollout_vjp_other_loss_matches_central_differences():
    """A loss the fused entry point does not offer, differentiated through
    the node: L = sum T_T^2 + 0.5 sum u_T^2 of the final state, FP64.
    The VJP for controls, beta and the initial u and density against FP64
    central differences (eps 1e-7) along 4 seeded directions, rel 1e-5."""
    import aether_fluid as af
    _announce("rollout vjp, loss sum T^2 + 0.5 sum u^2, fp64 16^3")
    n = 16
    s, pb = _rollout_setup(n, "fp64", tol=1e-13, iters=1000)
    rng = np.random.default_rng(11)
    init = {"u": rng.uniform(-0.5, 0.5, n ** 3), "density": rng.uniform(0, 1, n ** 3)}
    for k, a in init.items():
        s.upload(k, a)
    tape = af.SmokeRollout(s, nsteps=6, dt=0.02, patch_lo=pb["patch_lo"],
                           patch_hi=pb["patch_hi"], checkpoint_every=2)

    def loss_of(final):
        return float(np.sum(final["temperature"] ** 2) + 0.5 * np.sum(final["u"] ** 2))

    final = tape.forward(pb["controls"], 1.0)
    gc, gb, gi = tape.backward({"temperature": 2.0 * final["temperature"], "u": final["u"]})
    eps = 1e-7
    for q, (dc, db) in enumerate(_directions(4, len(pb["controls"]), seed=21)):
        du = rng.uniform(-1, 1, n ** 3)
        dr = rng.uniform(-1, 1, n ** 3)
        vals = []
        for sign in (1.0, -1.0):
            st = {"u": init["u"] + sign * eps * du, "density": init["density"] + sign * eps * dr}
            vals.append(loss_of(tape.forward(pb["controls"] + sign * eps * dc,
                                             1.0 + sign * eps * db, st)))
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = float(gc @ dc + gb * db + gi["u"] @ du + gi["density"] @ dr)
        print("vjp dir", q, fd, ad, abs(ad - fd) / abs(fd), flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd), (q, ad, fd)


def test_torch_function_over_the_node():
    """torch.autograd.Function form: the final state is a (5, N) tensor, a
    torch loss on it (here sum (rho_T - target)^2 + sum T_T^2) backpropagates
    through the library's adjoint into the controls, beta and the initial
    state, and equals the node's VJP with the same adjoints."""
    import torch
    import aether_fluid as af
    from aether_fluid.adjoint import torch_function
    _announce("torch function over the rollout node fp64 16^3")
    n = 16
    s, pb = _rollout_setup(n, "fp64")
    kw = dict(nsteps=4, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
              checkpoint_every=2)
    tape = af.SmokeRollout(s, **kw)
    final = tape.forward(pb["controls"], 1.0)
    lam = {"density": 2.0 * (final["density"] - pb["target"]),
           "temperature": 2.0 * final["temperature"]}
    gc, gb, gi = tape.backward(lam)
    fn = torch_function()
    c = torch.tensor(pb["controls"], dtype=torch.float64, requires_grad=True)
    b = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    init = torch.zeros((5, n ** 3), dtype=torch.float64, requires_grad=True)
    out = fn.apply(c, b, init, tape)
    tgt = torch.from_numpy(pb["target"])
    loss = ((out[3] - tgt) ** 2).sum() + (out[4] ** 2).sum()
    loss.backward()
    np.testing.assert_array_equal(out[3].detach().numpy(), final["density"])
    np.testing.assert_array_equal(c.grad.numpy(), gc)
    assert float(b.grad) == gb
    np.testing.assert_array_equal(init.grad[3].numpy(), gi["density"])


def test_adjoint_fp64_rollout_64_real_size_matches_central_differences():
    """The differentiable-rollout workload at its real size: 64^3, 40 steps
    of dt 0.02, checkpoint every 8, inflow-patch controls (32x4x32) and the
    buoyancy coefficient.  The FP64 adjoint against FP64 central differences
    at eps = 1e-7 (small enough not to straddle the clamp and floor switches
    of the discrete forward) along all 16 seeded directions, relative error
    <= 1e-5; the FP32 adjoint (FP64 reductions, fixed-point scatters)
    against the same differences, relative error <= 1e-4 in every
    direction (also with the adjoint's projections stopped at 1e-5, the
    benchmark's setting), and against the FP64 adjoint, whole-gradient
    relative L2 <= 1e-3.

    The FP32 rollout projects to relative residual 1e-6 (the workload's
    tight tolerance; FP32 CG attains a true residual of ~6e-7 here).  The
    gradient is the exact derivative of a piecewise-smooth forward (clamp
    arg-min/max, floor, near-wall selections): a trajectory that differs by
    ~1e-8 can sit on the other side of a selection in some cell, which moves
    the gradient by ~1e-4 relative.  Measured: FP32 at tolerance 1e-6 is
    within 9.8e-6 of the differences in every direction; an FP64 forward
    projected only to 1e-5 is 8.6e-5 (relative L2) off the tight FP64
    gradient, as is FP32 at 1e-5 or at 1e-7."""
    _announce("adjoint fp64/fp32 vs FD, rollout 64^3 x 40 steps, 16 directions")
    n = 64
    s64, pb = _rollout_setup(n, "fp64", tol=1e-12, iters=3000)
    s32, _ = _rollout_setup(n, "fp32", tol=1e-6, iters=1000)
    args = dict(nsteps=40, dt=0.02, patch_lo=pb["patch_lo"], patch_hi=pb["patch_hi"],
                target=pb["target"], checkpoint_every=8)
    assert len(pb["controls"]) == 32 * 4 * 32
    loss, gc, gb = s64.rollout(controls=pb["controls"], beta=1.0, **args)
    _, gc32, gb32 = s32.rollout(controls=pb["controls"], beta=1.0, **args)
    # the benchmark's setting: adjoint projections to 1e-5 only
    _, gc32a, gb32a = s32.rollout(controls=pb["controls"], beta=1.0, adjoint_cg_rel_tol=1e-5,
                                  **args)
    assert loss > 0 and np.all(np.isfinite(gc))
    assert np.linalg.norm(gc32 - gc) <= 1e-3 * np.linalg.norm(gc), np.linalg.norm(gc32 - gc)
    assert abs(gb32 - gb) <= 1e-3 * abs(gb)
    eps = 1e-7
    e64, e32, e32a = [], [], []
    for dc, db in _directions(16, len(pb["controls"]), seed=0xAE7E0003):
        lp, _, _ = s64.rollout(controls=pb["controls"] + eps * dc, beta=1.0 + eps * db,
                               grad=False, **args)
        lm, _, _ = s64.rollout(controls=pb["controls"] - eps * dc, beta=1.0 - eps * db,
                               grad=False, **args)
        fd = (lp - lm) / (2 * eps)
        e64.append(abs(float(gc @ dc + gb * db) - fd) / abs(fd))
        e32.append(abs(float(gc32 @ dc + gb32 * db) - fd) / abs(fd))
        e32a.append(abs(float(gc32a @ dc + gb32a * db) - fd) / abs(fd))
        print(f"64^3 dir fd {fd:.9e} fp64 {e64[-1]:.2e} fp32 {e32[-1]:.2e} "
              f"(adjoint 1e-5: {e32a[-1]:.2e})", flush=True)
    print("max fp64", max(e64), "max fp32", max(e32), "median fp32", np.median(e32),
          "max fp32 adjoint 1e-5", max(e32a), flush=True)
    assert max(e64) <= 1e-5, e64
    assert max(e32) <= 1e-4, e32
    assert max(e32a) <= 1e-4, e32a


def _volume_rays(n, seed):
    """n seeded rays for the volumetric samplers: origins around and inside
    the box [-1, 1]^3, directions normalised as Ray::new does (the oracle's
    Vec3::normalize), with edge cases: axis-parallel directions (infinite
    inv_dir), origins on a slab plane (0 * inf), far limits before the entry
    and t_min past the box."""
    from oracle.oracle import OracleFluid
    rng = np.random.default_rng(seed)
    org = rng.uniform(-2.0, 2.0, (n, 3))
    d = rng.normal(size=(n, 3))
    d[0:200, 1:] = 0.0          # along +-x
    d[200:400, :2] = 0.0        # along +-z
    org[400:600, 0] = -1.0      # on the x = -1 plane
    d[400:600, 0] = 0.0         # ... and parallel to it
    org[600:800] = rng.uniform(-0.9, 0.9, (200, 3))  # inside
    d = OracleFluid.normalize(d)
    t_min = rng.uniform(0.0, 0.05, n)
    t_min[800:900] = 50.0       # past the box
    t_far = rng.uniform(0.5, 6.0, n)
    t_far[900:1000] = 1e-3      # before the entry
    lam = rng.uniform(380.0, 750.0, (n, 4))
    return org, d, t_min, t_far, lam


def _volume_pair(precision, empty=False):
    import aether_fluid as af
    from oracle.oracle import OracleFluid
    dims = (24, 20, 28)
    rng = np.random.default_rng(9)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]),
                          indexing="ij")
    dens = (0.7 + 0.3 * np.sin(0.4 * i) * np.cos(0.3 * j) * np.cos(0.25 * k)
            + 0.1 * rng.random(i.shape)).ravel()
    temp = (1.2 * np.exp(-((i - 12) ** 2 + (j - 6) ** 2 + (k - 14) ** 2) / 40.0)).ravel()
    if empty:
        dens[:] = 0.0
    s = af.FluidSolver3D(*dims, precision=precision)
    s.upload("density", dens)
    s.upload("temperature", temp)
    o = OracleFluid(*dims)
    # the medium snapshots the solver's storage: the oracle gets those values
    o.field("density")[...] = s.download("density").reshape(dims[2], dims[1], dims[0])
    o.field("temperature")[...] = s.download("temperature").reshape(dims[2], dims[1], dims[0])
    return s, o


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_volumetric_sampling_vs_oracle(precision):
    """FluidMedium's path-tracer queries on the device (src/fluid.rs:793-910)
    against the oracle's restatement over 10^5 seeded rays, each with its own
    Xoshiro256++ state: delta tracking (hit or None, t, point, albedo,
    emission), ratio tracking, and HG phase sampling and pdf.  The final RNG
    states are equal, so every ray made the reference's draws; values agree
    to the device/host libm (log, exp, sin, cos) rounding, rtol 1e-12."""
    import aether_fluid as af
    from oracle.oracle import OracleFluid
    _announce(f"volumetric sampling vs oracle ({precision}), 1e5 rays")
    s, o = _volume_pair(precision)
    bounds = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    scale, g = 2.5, 0.55
    med = af.FluidMedium.from_solver(s, bounds, scale, g)
    assert med.max_density == o.max_density()
    n = 100000
    org, d, t_min, t_far, lam = _volume_rays(n, 17)
    rays = af.FluidMedium.make_rays(org, d, t_min, t_far)
    from oracle.oracle import splitmix64_stream
    st0 = np.array(splitmix64_stream(0xAE7E0005, 4 * n), dtype=np.uint64).reshape(n, 4)

    st_g, st_o = st0.copy(), st0.copy()
    got = med.sample_interaction(rays, lam, st_g)
    exp = o.medium_interaction(*bounds, scale, g, org, d, t_min, t_far, lam, st_o)
    np.testing.assert_array_equal(st_g, st_o)
    np.testing.assert_array_equal(got["hit"], exp[:, 0].astype(np.int32))
    h = got["hit"] == 1
    assert 0.05 < h.mean() < 0.95, h.mean()  # both outcomes well represented
    np.testing.assert_allclose(got["t"][h], exp[h, 1], rtol=1e-12)
    np.testing.assert_allclose(got["point"][h], exp[h, 2:5], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got["albedo"][h], exp[h, 5:9], rtol=1e-12)
    np.testing.assert_allclose(got["emission"][h], exp[h, 9:13], rtol=1e-12, atol=1e-300)
    assert (got["emission"][h] > 0).any() and (got["emission"][h] == 0).any()

    st_g, st_o = st0.copy(), st0.copy()
    tg = med.transmittance(rays, st_g)
    to = o.medium_transmittance(*bounds, scale, g, org, d, t_min, t_far, st_o)
    np.testing.assert_array_equal(st_g, st_o)
    np.testing.assert_allclose(tg, to, rtol=1e-12, atol=1e-300)
    assert (tg == 0).any() and (tg == 1).any() and ((tg > 0) & (tg < 1)).any()

    wo = OracleFluid.normalize(np.random.default_rng(4).normal(size=(n, 3)))
    for gg in (g, 0.0):
        med_g = af.FluidMedium.from_solver(s, bounds, scale, gg)
        st_g, st_o = st0.copy(), st0.copy()
        wi_g, pdf_g = med_g.sample_phase_hg(wo, st_g)
        wi_o, pdf_o = OracleFluid.phase_sample(gg, wo, st_o)
        np.testing.assert_array_equal(st_g, st_o)
        np.testing.assert_allclose(wi_g, wi_o, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(pdf_g, pdf_o, rtol=1e-12)
        np.testing.assert_allclose(med_g.phase_hg_pdf(wo, wi_o),
                                   OracleFluid.phase_pdf(gg, wo, wi_o), rtol=1e-14)
        med_g.close()
    med.close()


def test_volumetric_sampling_empty_medium_draws_nothing():
    """max_density <= 1e-6 (src/fluid.rs:799, 842): no interaction,
    transmittance 1, and no RNG draw."""
    import aether_fluid as af
    _announce("volumetric sampling, empty medium")
    s, o = _volume_pair("fp64", empty=True)
    med = af.FluidMedium.from_solver(s, ((-1, -1, -1), (1, 1, 1)), 2.5, 0.5)
    n = 1000
    org, d, t_min, t_far, lam = _volume_rays(n, 3)
    rays = af.FluidMedium.make_rays(org, d, t_min, t_far)
    st = np.arange(4 * n, dtype=np.uint64).reshape(n, 4) + 1
    before = st.copy()
    assert not med.sample_interaction(rays, lam, st)["hit"].any()
    np.testing.assert_array_equal(med.transmittance(rays, st), 1.0)
    np.testing.assert_array_equal(st, before)
