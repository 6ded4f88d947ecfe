"""CPU tier: pins of the f64 oracle (oracle/fluid_oracle.c).

Each test checks the oracle against something it does not restate: the
reference's own known answer and ratio tests, published RNG vectors, or a
property fixed by the mathematics (exactness on affine fields, integer
shifts, an independently assembled graph Laplacian, a dense least-squares
solve, the post-projection residual identity, closed-form telemetry), or
the first second of the reference's rendered soundtrack, which the first
nine steps of its default run drive.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleFluid, OracleRng, soundtrack, splitmix64_stream


def _interior(nx, ny, nz):
    m = np.zeros((nz, ny, nx), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


# ---------------------------------------------------------------- reference pins
def test_readme_known_answer():
    """README.md:54: verify.rs:229-250 prints 2.5751e0 -> 6.7289e-3 (99.7%)."""
    s = OracleFluid(16, 20, 16)
    s.solver_kind, s.solver_iters = 0, 40
    s.inject_sources(0.5, 0.1)
    s.apply_buoyancy_and_vorticity_confinement(0.1)
    pre = s.compute_divergence_norm()
    s.project()
    post = s.compute_divergence_norm()
    assert f"{pre:.4e}" == "2.5751e+00"
    assert f"{post:.4e}" == "6.7289e-03"
    assert f"{(1 - post / pre) * 100:.1f}" == "99.7"
    assert post <= 0.03 * pre  # verify.rs:239


@pytest.mark.parametrize("kind,iters", [(0, 45), (1, 90)])
def test_fluid_rs_projection_ratio(kind, iters):
    """src/fluid.rs:918-956."""
    s = OracleFluid(18, 22, 18)
    s.solver_kind, s.solver_iters = kind, iters
    u, v, w = s.field("u"), s.field("v"), s.field("w")
    for k in range(6, 12):
        for j in range(5, 11):
            for i in range(6, 12):
                u[k, j, i] = (i - 8.5) * 3.0
                v[k, j, i] = (j - 7.5) * 4.5
                w[k, j, i] = (k - 8.5) * 2.5
    s.enforce_wall_velocity_bc()
    before = s.compute_divergence_norm()
    assert before > 1.0
    s.project()
    assert s.compute_divergence_norm() < before * 0.02


def test_fluid_rs_positivity_two_steps():
    """src/fluid.rs:959-968."""
    s = OracleFluid(16, 20, 16)
    f1 = s.step(0.08, 0.2)
    f2 = s.step(0.08, 0.4)
    assert f1[1] > 0.0 and f2[2] > 0.0
    d = s.field("density")
    assert np.all(np.isfinite(d)) and np.all(d >= 0.0)


# ---------------------------------------------------------------- RNG
def test_xoshiro256pp_published_vector():
    """Xoshiro256++ from state {1,2,3,4}: the published first ten outputs."""
    r = OracleRng(state=[1, 2, 3, 4])
    got = [r.next_u64() for _ in range(10)]
    assert got == [41943041, 58720359, 3588806011781223, 3591011842654386,
                   9228616714210784205, 9973669472204895162, 14011001112246962877,
                   12406186145184390807, 15849039046786891736, 10450023813501588000]


def test_splitmix64_published_vector():
    assert splitmix64_stream(0, 3)[0] == 0xE220A8397B1DCDAF


def test_next_f64_scaling():
    # next_f64 = (x >> 11) * 2^-53: for state {1,2,3,4} the first draw is 41943041 >> 11
    r = OracleRng(state=[1, 2, 3, 4])
    assert r.next_f64() == (41943041 >> 11) * 2.0 ** -53


def test_product_rng_matches_oracle_rng():
    import aether_fluid as af
    vals, st = af.rng_fill(0xAE7E0001, -1.0, 1.0, 1000)
    r = OracleRng(0xAE7E0001)
    exp = np.array([r.range_f64(-1.0, 1.0) for _ in range(1000)])
    np.testing.assert_array_equal(vals, exp)
    more, _ = af.rng_fill(st, 0.0, 1.0, 10)
    np.testing.assert_array_equal(more, [r.next_f64() for _ in range(10)])


# ---------------------------------------------------------------- sampling / advection
def test_trilinear_exact_on_affine_field():
    s = OracleFluid(9, 11, 13)
    k, j, i = np.meshgrid(np.arange(13), np.arange(11), np.arange(9), indexing="ij")
    s.field("density")[...] = 0.25 + 1.5 * i - 0.75 * j + 2.0 * k
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = rng.uniform(0.5, 9 - 1.5), rng.uniform(0.5, 11 - 1.5), rng.uniform(0.5, 13 - 1.5)
        exp = 0.25 + 1.5 * x - 0.75 * y + 2.0 * z
        assert abs(s.sample_trilinear("density", x, y, z) - exp) < 1e-12


def test_trilinear_clamps_outside():
    s = OracleFluid(8, 8, 8)
    f = s.field("density")
    f[...] = np.random.default_rng(1).uniform(size=f.shape)
    a = s.sample_trilinear("density", -5.0, 3.2, 100.0)
    b = s.sample_trilinear("density", 0.5, 3.2, 8 - 1.500001)
    assert a == b


# ---------------------------------------------------------------- FluidMedium
def _feather(u):
    """fluid.rs:755-758 restated independently (numpy clip)."""
    fx = np.clip(4.0 * u[0] * (1.0 - u[0]), 0.0, 1.0)
    fz = np.clip(4.0 * u[2] * (1.0 - u[2]), 0.0, 1.0)
    fy = np.clip(1.0 - max(u[1] - 0.95, 0.0) / 0.05, 0.0, 1.0)
    return math.sqrt(fx * fz * fy)


def test_medium_sample_affine_fields_and_feather():
    """Inside the clamp box, trilinear interpolation of affine density and
    temperature fields is exact; the result is scaled by the edge feather
    and sampled at u * (n - 1) of the bounds (fluid.rs:739-788)."""
    nx, ny, nz = 10, 12, 9
    s = OracleFluid(nx, ny, nz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    s.field("density")[...] = 0.5 + 0.25 * i + 0.125 * j - 0.0625 * k
    s.field("temperature")[...] = 2.0 - 0.5 * i + 0.75 * j + 0.25 * k
    lo, hi = (-1.0, 0.0, 2.0), (3.0, 2.5, 4.0)
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.uniform(lo[a], hi[a], 300) for a in range(3)])
    got = s.medium_sample(lo, hi, pts)
    for p, (d, t) in zip(pts, got):
        u = [(p[a] - lo[a]) / (hi[a] - lo[a]) for a in range(3)]
        g = [min(max(u[a] * (n - 1), 0.5), n - 1.500001) for a, n in enumerate((nx, ny, nz))]
        f = _feather(u)
        assert abs(d - (0.5 + 0.25 * g[0] + 0.125 * g[1] - 0.0625 * g[2]) * f) < 1e-12
        assert abs(t - (2.0 - 0.5 * g[0] + 0.75 * g[1] + 0.25 * g[2]) * f) < 1e-12


def test_medium_sample_outside_and_feather_values():
    s = OracleFluid(8, 8, 8)
    s.field("density")[...] = 1.0
    s.field("temperature")[...] = 3.0
    lo, hi = (0.0, 0.0, 0.0), (1.0, 2.0, 1.0)
    out = s.medium_sample(lo, hi, [(-0.01, 1.0, 0.5), (0.5, 2.01, 0.5), (0.5, 1.0, 1.5)])
    assert np.all(out == 0.0)
    # centre: feather 1; uy = 0.975: fy = 0.5, feather sqrt(0.5); x wall: 0
    out = s.medium_sample(lo, hi, [(0.5, 1.0, 0.5), (0.5, 1.95, 0.5), (0.0, 1.0, 0.5)])
    assert out[0, 0] == 1.0 and out[0, 1] == 3.0
    assert abs(out[1, 0] - math.sqrt(0.5)) < 1e-15
    assert out[2, 0] == 0.0 and out[2, 1] == 0.0


def test_medium_max_density_starts_from_zero():
    """fluid.rs:704-707: max over density starting from 0."""
    s = OracleFluid(6, 6, 6)
    s.field("density")[...] = -1.0
    assert s.max_density() == 0.0
    s.field("density")[2, 3, 4] = 7.5
    assert s.max_density() == 7.5


def test_maccormack_integer_shift_is_exact():
    """Uniform velocity moving exactly one cell per step: interior cells away
    from the walls receive the upwind neighbour's value times dissipation."""
    nx = ny = nz = 14
    s = OracleFluid(nx, ny, nz, h=0.1)
    dt = 0.1  # dt/h * 1.0 = one cell
    s.field("u")[...] = 1.0
    s.field("v")[...] = 0.0
    s.field("w")[...] = 0.0
    rho = np.random.default_rng(2).uniform(size=(nz, ny, nx))
    s.field("density")[...] = rho
    s.advect_maccormack_in_place(3, dt, 0.9)
    out = s.field("density")
    # back-trace x = i - 1 must stay >= 1.5 and forward trace i + 1 <= n - 2.5
    sl = (slice(3, nz - 3), slice(3, ny - 3), slice(3, nx - 3))
    np.testing.assert_array_equal(out[sl], 0.9 * rho[sl[0], sl[1], slice(2, nx - 4)])


def test_maccormack_constant_field():
    s = OracleFluid(10, 12, 9)
    rng = np.random.default_rng(3)
    for n in ("u", "v", "w"):
        s.field(n)[...] = rng.uniform(-1, 1, size=(9, 12, 10))
    s.field("temperature")[...] = 0.7
    s.advect_maccormack_in_place(4, 0.05, 0.985)
    t = s.field("temperature")
    inside = _interior(10, 12, 9)
    np.testing.assert_allclose(t[inside], 0.7 * 0.985, rtol=1e-15)
    assert np.all(t[~inside] == 0.0)


def test_maccormack_stays_within_clamp_and_nonnegative():
    s = OracleFluid(16, 16, 16)
    rng = np.random.default_rng(4)
    for n in ("u", "v", "w"):
        s.field(n)[...] = rng.uniform(-3, 3, size=(16, 16, 16))
    rho = rng.uniform(0, 1, size=(16, 16, 16))
    s.field("density")[...] = rho
    s.advect_maccormack_in_place(3, 0.1, 1.0)
    d = s.field("density")
    assert d.min() >= 0.0 and d.max() <= rho.max()


# ---------------------------------------------------------------- Poisson operator
def _dense_operator(s):
    n = s.n
    A = np.zeros((n, n))
    e = np.zeros(n)
    for c in range(n):
        e[:] = 0.0
        e[c] = 1.0
        A[:, c] = s.apply_laplacian(e).ravel()
    return A


def _graph_laplacian(nx, ny, nz):
    """L = D - Adj of the graph whose nodes are interior cells and whose edges
    join face neighbours (the pure-Neumann 7-point Laplacian)."""
    n = nx * ny * nz
    L = np.zeros((n, n))
    idx = lambda i, j, k: (k * ny + j) * nx + i  # noqa: E731
    for k in range(1, nz - 1):
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                a = idx(i, j, k)
                for di, dj, dk in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 1 <= ii <= nx - 2 and 1 <= jj <= ny - 2 and 1 <= kk <= nz - 2:
                        b = idx(ii, jj, kk)
                        L[a, a] += 1
                        L[b, b] += 1
                        L[a, b] -= 1
                        L[b, a] -= 1
    return L


def test_laplacian_is_interior_graph_laplacian():
    s = OracleFluid(6, 7, 8)
    A = _dense_operator(s)
    np.testing.assert_array_equal(A, _graph_laplacian(6, 7, 8))
    assert np.allclose(A, A.T)
    ones = _interior(6, 7, 8).ravel().astype(float)
    assert np.abs(A @ ones).max() == 0.0
    assert np.linalg.eigvalsh(A).min() > -1e-12


def test_cg_matches_dense_least_squares():
    nx, ny, nz = 7, 6, 8
    s = OracleFluid(nx, ny, nz)
    inside = _interior(nx, ny, nz).ravel()
    rng = np.random.default_rng(5)
    b = np.zeros(s.n)
    b[inside] = rng.normal(size=inside.sum())
    b[inside] -= b[inside].mean()  # compatible with the constant null space
    s.field("divergence")[...] = b.reshape(nz, ny, nx)
    s.solve_poisson_conjugate_gradient(500)
    p = s.field("pressure").ravel()[inside]
    L = _graph_laplacian(nx, ny, nz)[np.ix_(inside, inside)]
    ref, *_ = np.linalg.lstsq(L, b[inside], rcond=None)
    np.testing.assert_allclose(p - p.mean(), ref - ref.mean(), atol=1e-8)
    st = s.stats()
    assert st["converged"] and st["r2"] < 1e-20


def test_cg_zero_rhs_no_iterations():
    s = OracleFluid(10, 10, 10)
    s.solver_iters = 50
    s.project()
    assert s.stats()["iters"] == 0


def test_cg_rel_tol_stops_early():
    s = OracleFluid(16, 16, 16)
    rng = np.random.default_rng(6)
    for n in ("u", "v", "w"):
        s.field(n)[...] = rng.uniform(-1, 1, size=(16, 16, 16))
    s.solver_iters, s.cg_rel_tol = 500, 1e-4
    s.project()
    st = s.stats()
    assert st["converged"] and st["r2"] <= 1e-8 * st["b2"] and 0 < st["iters"] < 500


def test_projection_residual_identity():
    """After project(), -h * div(u) equals the CG residual cell by cell:
    D(u - G p/h) = D u + A p / h, so -h D u_new = b - A p = r."""
    s = OracleFluid(12, 14, 10)
    rng = np.random.default_rng(7)
    for n in ("u", "v", "w"):
        s.field(n)[...] = rng.uniform(-1, 1, size=(10, 14, 12))
    s.solver_iters = 15  # stop early so that r is far from zero
    s.project()
    r = s.field("cg_r").copy()
    s.compute_divergence_norm()  # stores -h*div(u) of the projected field
    b_new = s.field("divergence")
    assert np.abs(r).max() > 1e-3
    np.testing.assert_allclose(b_new, r, atol=1e-12)


def test_rbgs_parallel_equals_sequential():
    """Red-black colouring: the OpenMP build (parallel over k within a
    colour) gives the sequential result bit for bit."""
    outs = []
    for omp in (False, True):
        s = OracleFluid(14, 12, 16, omp=omp)
        rng = np.random.default_rng(8)
        for n in ("u", "v", "w"):
            s.field(n)[...] = rng.uniform(-1, 1, size=(16, 12, 14))
        s.solver_kind, s.solver_iters = 1, 25
        s.project()
        outs.append(s.field("pressure").copy())
    np.testing.assert_array_equal(outs[0], outs[1])


# ---------------------------------------------------------------- forces
def test_linear_velocity_has_no_confinement():
    """Curl of a linear field is constant, so grad|w| = 0 and the
    confinement force vanishes; only buoyancy and the wall BC act."""
    nx, ny, nz = 10, 11, 12
    s = OracleFluid(nx, ny, nz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    # dyadic coefficients: every difference is exact, so |w| is bit-identical
    # in all cells and the 1e-10 guard of src/fluid.rs:298 makes N = 0
    s.field("u")[...] = 0.25 * j + 0.125 * k
    s.field("v")[...] = -0.5 * k + 0.0625 * i
    s.field("w")[...] = 0.375 * i
    rng = np.random.default_rng(9)
    T = rng.uniform(size=(nz, ny, nx))
    R = rng.uniform(size=(nz, ny, nx))
    s.field("temperature")[...] = T
    s.field("density")[...] = R
    u0, v0, w0 = (s.field(n).copy() for n in ("u", "v", "w"))
    s.vorticity_eps, s.buoyancy_beta, s.gravity_alpha = 50.0, 2.0, 0.5
    dt = 0.1
    s.apply_buoyancy_and_vorticity_confinement(dt)
    inside = _interior(nx, ny, nz)
    ev = v0 + np.where(inside, dt * (2.0 * T - 0.5 * R), 0.0)
    # wall masks of src/fluid.rs:241-249
    mu = np.ones_like(inside)
    mu[:, :, :2] = mu[:, :, -1:] = mu[:, :1, :] = mu[:, -1:, :] = mu[:1] = mu[-1:] = False
    mv = np.ones_like(inside)
    mv[:, :2, :] = mv[:, -1:, :] = mv[:, :, :1] = mv[:, :, -1:] = mv[:1] = mv[-1:] = False
    mw = np.ones_like(inside)
    mw[:2] = mw[-1:] = mw[:, :, :1] = mw[:, :, -1:] = mw[:, :1, :] = mw[:, -1:, :] = False
    np.testing.assert_allclose(s.field("u"), np.where(mu, u0, 0.0), atol=1e-13)
    np.testing.assert_allclose(s.field("v"), np.where(mv, ev, 0.0), atol=1e-13)
    np.testing.assert_allclose(s.field("w"), np.where(mw, w0, 0.0), atol=1e-13)


def test_sphere_source_sets_values_inside_only():
    s = OracleFluid(20, 20, 20)
    s.set_source(1, (10.0, 3.0, 9.5), 4.0, 1.0, 2.0)
    s.inject_sources(0.0, 0.1)
    k, j, i = np.meshgrid(np.arange(20), np.arange(20), np.arange(20), indexing="ij")
    inside = ((i - 10.0) ** 2 + (j - 3.0) ** 2 + (k - 9.5) ** 2 < 16.0) & _interior(20, 20, 20)
    np.testing.assert_array_equal(s.field("density"), np.where(inside, 1.0, 0.0))
    np.testing.assert_array_equal(s.field("temperature"), np.where(inside, 2.0, 0.0))


def test_jets_at_time_zero_closed_form():
    """At t = 0: pulse = 0.65 + 0.35 sin 0 + 0.25 max(cos 0, 0) = 0.9, angle
    = 0, so jet 1 sits at (cx + 0.14 nx, base_j + 1.5, cz) pushing
    (0, 7.2*0.9, 2.6) and jet 2 at (cx - 0.14 nx, base_j + 2, cz) pushing
    (0, 6.4*0.9, -2.6) (src/fluid.rs:538-560)."""
    nx, ny, nz = 40, 30, 40
    s = OracleFluid(nx, ny, nz)
    dt = 0.1
    s.inject_sources(0.0, dt)
    r = nx * 0.105
    jets = [((20 + 0.14 * nx, 4.5, 20.0), (0.0, 7.2 * 0.9, 2.6)),
            ((20 - 0.14 * nx, 5.0, 20.0), (0.0, 6.4 * 0.9, -2.6))]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rho = np.zeros((nz, ny, nx))
    vz = np.zeros((nz, ny, nx))
    for (cx, cy, cz), vel in jets:
        d2 = (i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2
        wf = np.where(d2 < r * r, np.exp(-2.2 * d2 / (r * r)), 0.0)
        rho = np.minimum(rho + dt * 9.5 * 0.9 * wf, 3.8)
        vz = vz + dt * vel[2] * wf * 4.2
    np.testing.assert_allclose(s.field("density"), rho, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(s.field("w"), vz, rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------- telemetry
def test_telemetry_uniform_velocity_and_unit_mass():
    nx, ny, nz = 9, 8, 10
    s = OracleFluid(nx, ny, nz)
    h = s.h
    s.field("u")[...] = 0.3
    s.field("v")[...] = -0.4
    s.field("w")[...] = 1.2
    s.field("density")[4, 3, 6] = 2.0  # (i, j, k) = (6, 3, 4)
    t = s.compute_telemetry(1.25)
    interior_n = (nx - 2) * (ny - 2) * (nz - 2)
    assert t[0] == 1.25
    assert math.isclose(t[1], 0.5 * (0.09 + 0.16 + 1.44) * h ** 3 * interior_n, rel_tol=1e-12)
    assert t[2] == 0.0 and t[3] == 0.0
    assert math.isclose(t[4], 2.0 * h ** 3, rel_tol=1e-15)
    exp = [2 * 6 / (nx - 1) - 1, 2 * 3 / (ny - 1) - 1, 2 * 4 / (nz - 1) - 1]
    np.testing.assert_allclose(t[5:8], exp, rtol=1e-14)


def test_telemetry_solid_body_rotation_enstrophy():
    """u = -c*(y), v = c*(x) (about a vertical axis): curl = (0, 0, 2c) in
    index units /h... with central differences exact on linear fields."""
    nx, ny, nz = 10, 10, 10
    s = OracleFluid(nx, ny, nz)
    h = s.h
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = 0.7
    s.field("u")[...] = -c * j * h
    s.field("v")[...] = c * i * h
    t = s.compute_telemetry(0.0)
    interior_n = (nx - 2) * (ny - 2) * (nz - 2)
    assert math.isclose(t[3], 2 * c, rel_tol=1e-12)
    assert math.isclose(t[2], 0.5 * (2 * c) ** 2 * h ** 3 * interior_n, rel_tol=1e-12)


# ------------------------------------------------------- soundtrack head pin
def _soundtrack_fixture():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "soundtrack_head.npz")
    with np.load(path) as z:
        return {k: z[k] for k in z.files}


def _main_rs_frames(eps=6.5, iters=28, beta=None, nframes=11):
    """Telemetry of the first steps of the main.rs:82-94 run (36x52x36,
    dt 0.08, timeline 15 * i / 119), padded to 120 rows with the last one."""
    s = OracleFluid(36, 52, 36)
    s.vorticity_eps, s.solver_iters = eps, iters
    if beta is not None:
        s.buoyancy_beta = beta
    fr = np.array([s.step(0.08, 15.0 * (i / 119)) for i in range(nframes)])
    return np.concatenate([fr, np.repeat(fr[-1:], 120 - nframes, 0)])


def _head_diff(frames, fx):
    n, st = int(fx["nsamples"]), int(fx["stride"])
    pcm = soundtrack(frames, fx["norm"][0], fx["norm"][1], n)[::st].astype(np.int64)
    return pcm - fx["pcm_head"].astype(np.int64)


def test_soundtrack_head_reproduces_reference_wav():
    """renders/simulation_soundtrack.wav, first second: the oracle's first
    steps of the default run (frames 0-8 are read), through the audio.rs:397-639 restatement
    and the two fitted normalisers, give every stored sample within 1 LSB of
    24-bit (platform libm rounding at the dither quantiser)."""
    fx = _soundtrack_fixture()
    d = _head_diff(_main_rs_frames(), fx)
    assert np.abs(d).max() <= 1
    assert np.mean(d == 0) > 0.999


@pytest.mark.parametrize("knob", [dict(iters=27), dict(eps=6.45), dict(beta=6.79)])
def test_soundtrack_head_catches_planted_mistakes(knob):
    """One CG iteration fewer, or a 1% / 0.15% change of the confinement or
    buoyancy coefficient, moves the first second by far more than 1 LSB."""
    fx = _soundtrack_fixture()
    d = _head_diff(_main_rs_frames(**knob), fx)
    assert np.abs(d).max() > 10000


def test_soundtrack_first_second_reads_frames_0_to_8_only():
    """audio.rs:375-379: sample t reads frames floor(t / 15 * 119) and the
    next, so for t < 1 s (u < 7.94) frames past 8 cannot reach the first
    second; the synthesis is deterministic (fixed dither seed, audio.rs:623)."""
    fr = _main_rs_frames()
    alt = fr.copy()
    alt[9:, 1:] *= 3.0
    a = soundtrack(fr, 0.06, 190.0, 48000)
    np.testing.assert_array_equal(a, soundtrack(alt, 0.06, 190.0, 48000))
    np.testing.assert_array_equal(a, soundtrack(fr, 0.06, 190.0, 48000))
    alt[8, 1] *= 1.01
    assert not np.array_equal(a, soundtrack(alt, 0.06, 190.0, 48000))


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    """The checker itself under ASan + UBSan (oracle/sanitize_check.c): every
    public oracle function on small scenarios, including the README known
    answer; any out-of-bounds access, leak or undefined behaviour aborts."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-s", "-C", here, "sanitize"], check=True)
    res = subprocess.run([os.path.join(here, "oracle_sanitize_check")], capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-4000:]
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr
    assert "known answer 2.5751e+00 6.7289e-03" in res.stdout


# --- FluidMedium volumetric sampling (src/fluid.rs:793-910) -----------------
# The oracle restates delta tracking, ratio tracking and Henyey-Greenstein
# sampling; these checks pin it by what the mathematics fixes: ratio tracking
# is unbiased (E[T] = exp(-tau)), delta tracking's free flights follow
# P(no interaction before s) = exp(-tau(s)), HG samples follow the HG law
# (<cos> = g, its CDF), the pdf integrates to 1, and the spectral curves
# reach their channel values at the ends of the visible range.

def _vol_medium(seed=3):
    from oracle.oracle import OracleFluid
    o = OracleFluid(12, 10, 14)
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(14), np.arange(10), np.arange(12), indexing="ij")
    o.field("density")[...] = 0.6 + 0.3 * np.sin(0.7 * i) * np.cos(0.5 * j) + 0.05 * rng.random(
        (14, 10, 12))
    o.field("temperature")[...] = 0.8 * np.exp(-((i - 6) ** 2 + (k - 7) ** 2) / 18.0)
    return o


def _tau(o, bmin, bmax, scale, org, d, t0, t1, n=40001):
    """Optical depth along the ray by Simpson's rule on the pinned
    sample_density_and_temp."""
    ts = np.linspace(t0, t1, n)
    pts = org[None, :] + d[None, :] * ts[:, None]
    sig = o.medium_sample(bmin, bmax, pts)[:, 0] * scale
    w = np.ones(n)
    w[1:-1:2], w[2:-1:2] = 4, 2
    return (t1 - t0) / (3 * (n - 1)) * float(w @ sig)


def _seeded_states(n, seed):
    from oracle.oracle import splitmix64_stream
    return np.array(splitmix64_stream(seed, 4 * n), dtype=np.uint64).reshape(n, 4)


def test_oracle_ratio_tracking_is_unbiased():
    from oracle.oracle import OracleFluid
    o = _vol_medium()
    bmin, bmax, scale = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 1.5
    org = np.array([-1.5, -0.2, 0.1])
    d = OracleFluid.normalize(np.array([[1.0, 0.15, -0.1]]))[0]
    n = 40000
    states = _seeded_states(n, 11)
    far = 10.0
    tr = o.medium_transmittance(bmin, bmax, scale, 0.0, np.tile(org, (n, 1)), np.tile(d, (n, 1)),
                                np.zeros(n), np.full(n, far), states)
    # the ray enters at x = -1 and leaves at x = 1 (t = (x - ox) / dx)
    t0, t1 = (-1.0 - org[0]) / d[0], (1.0 - org[0]) / d[0]
    expect = np.exp(-_tau(o, bmin, bmax, scale, org, d, t0, t1))
    se = tr.std() / np.sqrt(n)
    assert 0.1 < expect < 0.9
    assert abs(tr.mean() - expect) < 5 * se, (tr.mean(), expect, se)
    assert tr.min() >= 0.0 and tr.max() <= 1.0
    # a ray that misses the box, and a medium without density, draw nothing
    st2 = _seeded_states(3, 12)
    before = st2.copy()
    miss = o.medium_transmittance(bmin, bmax, scale, 0.0, np.tile([5.0, 5.0, 5.0], (3, 1)),
                                  np.tile(d, (3, 1)), np.zeros(3), np.full(3, far), st2)
    np.testing.assert_array_equal(miss, 1.0)
    np.testing.assert_array_equal(st2, before)


def test_oracle_delta_tracking_free_flight_law():
    from oracle.oracle import OracleFluid
    o = _vol_medium()
    bmin, bmax, scale = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 1.5
    org = np.array([-1.5, -0.2, 0.1])
    d = OracleFluid.normalize(np.array([[1.0, 0.15, -0.1]]))[0]
    n = 40000
    states = _seeded_states(n, 21)
    lam = np.tile([400.0, 500.0, 600.0, 700.0], (n, 1))
    t0, t1 = (-1.0 - org[0]) / d[0], (1.0 - org[0]) / d[0]
    for surface in (10.0, 0.5 * (t0 + t1)):
        out = o.medium_interaction(bmin, bmax, scale, 0.0, np.tile(org, (n, 1)),
                                   np.tile(d, (n, 1)), np.zeros(n), np.full(n, surface), lam,
                                   states)
        hit = out[:, 0] == 1.0
        lim = min(t1, surface)
        p_none = np.exp(-_tau(o, bmin, bmax, scale, org, d, t0, lim))
        se = np.sqrt(p_none * (1 - p_none) / n)
        assert abs((~hit).mean() - p_none) < 5 * se, ((~hit).mean(), p_none)
        th = out[hit, 1]
        assert th.min() >= t0 and th.max() < lim
        np.testing.assert_allclose(out[hit, 2:5], org + d * th[:, None], rtol=0, atol=1e-12)
        # the CDF of the interaction distance at the middle of the segment
        s = 0.5 * (t0 + lim)
        p_before = 1.0 - np.exp(-_tau(o, bmin, bmax, scale, org, d, t0, s))
        se2 = np.sqrt(p_before * (1 - p_before) / n)
        assert abs((hit & (out[:, 1] < s)).mean() - p_before) < 5 * se2


def test_oracle_emission_and_albedo_curves():
    """Scatter albedo SmoothRgb(0.84, 0.90, 0.96): blue weight at 380 nm, red
    at 750 nm, green between; emission is 0 where the temperature is at most
    0.35 and peaks at the 486 nm line where the core is hot."""
    o = _vol_medium()
    o.field("temperature")[...] = 0.0
    o.field("temperature")[:, :, :6] = 1.0  # hot half x < 0 (glow (1 - 0.35)^2 0.95)
    bmin, bmax = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    n = 4000
    states = _seeded_states(n, 31)
    lam = np.tile([380.0, 540.0, 486.0, 750.0], (n, 1))
    out = o.medium_interaction(bmin, bmax, 50.0, 0.0, np.tile([-1.5, 0.0, 0.0], (n, 1)),
                               np.tile([1.0, 0.0, 0.0], (n, 1)), np.zeros(n), np.full(n, 10.0),
                               lam, states)
    hit = out[:, 0] == 1.0
    assert hit.mean() > 0.99
    alb = out[hit][0, 5:9]
    assert abs(alb[0] - 0.96) < 1e-3 and abs(alb[3] - 0.84) < 1e-3 and abs(alb[1] - 0.90) < 1e-2
    em = out[hit, 9:13]
    hot = em[:, 2] > 0
    assert hot.any() and (~hot).any()
    # at 486 nm the line (amp 1.6) dominates the continuum of the hot core
    assert np.all(em[hot, 2] > em[hot, 1])


def test_oracle_henyey_greenstein_law():
    from oracle.oracle import OracleFluid
    n = 60000
    rng = np.random.default_rng(5)
    wo = OracleFluid.normalize(rng.normal(size=(n, 3)))
    for g in (0.6, -0.3, 0.0):
        states = _seeded_states(n, 41)
        wi, pdf = OracleFluid.phase_sample(g, wo, states)
        np.testing.assert_allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-12)
        mu = np.einsum("ij,ij->i", -wo, wi)
        assert abs(mu.mean() - g) < 5 * mu.std() / np.sqrt(n), (g, mu.mean())
        np.testing.assert_allclose(pdf, OracleFluid.phase_pdf(g, wo, wi), rtol=1e-12)
        if g != 0.0:
            # HG CDF of cos(theta): (1-g^2)/(2g) [(1+g^2-2g mu)^-1/2 - (1+g)^-1]
            xs = np.sort(mu)
            cdf = (1 - g * g) / (2 * g) * ((1 + g * g - 2 * g * xs) ** -0.5 - 1 / (1 + g))
            ks = np.max(np.abs(cdf - np.arange(1, n + 1) / n))
            assert ks < 1.63 / np.sqrt(n), (g, ks)  # 1% level of the KS statistic
        else:
            np.testing.assert_allclose(pdf, 1 / (4 * np.pi))
    # the pdf integrates to 1 over the sphere (2 pi int p(mu) dmu)
    for g in (0.85, -0.5, 0.2):
        mus = np.linspace(-1, 1, 200001)
        w_o = np.tile([0.0, 0.0, -1.0], (len(mus), 1))  # -wo = +z
        w_i = np.stack([np.sqrt(1 - mus ** 2), np.zeros_like(mus), mus], axis=1)
        p = OracleFluid.phase_pdf(g, w_o, w_i)
        assert abs(2 * np.pi * np.trapezoid(p, mus) - 1.0) < 1e-6
