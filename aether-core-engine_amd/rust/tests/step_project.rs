//! The reference's fluid tests (src/fluid.rs:916-968) and its projection
//! self-check (src/verify.rs:229-250, README.md:54) through the wrapper: the
//! same calls a reference caller makes, on the device.  Needs a GPU and
//! libaether_fluid_hip.so (built by `make -C aether-core-engine_amd`).

use aether_fluid_hip::{FluidMedium, FluidSolver3D, PoissonSolverKind, SmokeTape};

#[test]
fn projection_cg_and_red_black_sor_remove_divergence() {
    for kind in [PoissonSolverKind::ConjugateGradient, PoissonSolverKind::RedBlackGaussSeidel] {
        let mut solver = FluidSolver3D::new(18, 22, 18);
        solver.solver_kind = kind;
        solver.solver_iters = if kind == PoissonSolverKind::ConjugateGradient { 45 } else { 90 };
        {
            // a divergent burst on the 6x6x6 block, written through the
            // host mirrors of the device arrays
            let ids: Vec<(usize, usize, usize, usize)> = (6..12)
                .flat_map(|k| (5..11).flat_map(move |j| (6..12).map(move |i| (i, j, k))))
                .map(|(i, j, k)| (i, j, k, (k * 22 + j) * 18 + i))
                .collect();
            let mut u = solver.u_mut();
            for &(i, _, _, id) in &ids {
                u[id] = (i as f64 - 8.5) * 3.0;
            }
            drop(u);
            let mut v = solver.v_mut();
            for &(_, j, _, id) in &ids {
                v[id] = (j as f64 - 7.5) * 4.5;
            }
            drop(v);
            let mut w = solver.w_mut();
            for &(_, _, k, id) in &ids {
                w[id] = (k as f64 - 8.5) * 2.5;
            }
        }
        solver.enforce_wall_velocity_bc();
        let before = solver.compute_divergence_norm();
        assert!(before > 1.0);
        solver.project();
        let after = solver.compute_divergence_norm();
        assert!(after < before * 0.02, "{kind:?}: {before} -> {after}");
    }
}

#[test]
fn steps_keep_energy_and_positive_density() {
    let mut solver = FluidSolver3D::new(16, 20, 16);
    let f1 = solver.step(0.08, 0.2);
    let f2 = solver.step(0.08, 0.4);
    assert!(f1.kinetic_energy > 0.0);
    assert!(f2.enstrophy > 0.0);
    assert!(solver.density().iter().all(|&d| d >= 0.0 && d.is_finite()));
    assert!(!solver.nonfinite());
}

#[test]
fn verify_self_check_known_answer() {
    let mut solver = FluidSolver3D::new(16, 20, 16);
    solver.solver_iters = 40;
    solver.inject_cyberpunk_sources(0.5, 0.1);
    solver.apply_buoyancy_and_vorticity_confinement(0.1);
    let pre = solver.compute_divergence_norm();
    solver.project();
    let post = solver.compute_divergence_norm();
    assert_eq!(format!("{pre:.4e} -> {post:.4e}"), "2.5751e0 -> 6.7289e-3");
}

#[test]
fn medium_and_tape() {
    let mut solver = FluidSolver3D::new(16, 16, 16);
    let _ = solver.step(0.08, 0.0);
    let medium = FluidMedium::from_solver(&solver, [-1.0; 3], [1.0; 3], 2.0, 0.4);
    assert!(medium.max_density > 0.0);
    let (d, _t) = medium.sample_density_and_temp([0.0, -0.5, 0.0]);
    assert!(d >= 0.0);
    let mut rng = [1u64, 2, 3, 4];
    let tr = medium.transmittance(
        &aether_fluid_hip::fluid::Ray { origin: [-2.0, -0.5, 0.0], dir: [1.0, 0.0, 0.0], t_min: 0.0 },
        10.0,
        &mut rng,
    );
    assert!((0.0..=1.0).contains(&tr));

    let mut tape = SmokeTape::new(&mut solver, 4, 0.02, [4, 1, 4], [12, 2, 12], 2);
    let controls = vec![0.5; 8 * 8];
    let fin = tape.forward(&controls, 1.0);
    let lam = aether_fluid_hip::tape::RolloutState {
        density: fin.density.iter().map(|x| 2.0 * x).collect(),
        ..Default::default()
    };
    let g = tape.backward(&lam);
    assert_eq!(g.controls.len(), 64);
    assert!(g.controls.iter().all(|x| x.is_finite()));
}
