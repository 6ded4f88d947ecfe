//! Reverse-mode node of a differentiable smoke rollout, in the shape of a
//! `torch.autograd.Function`: `forward` runs the checkpointed forward sweep
//! on the device and records the tape there; `backward` takes the adjoint
//! of the final state (any loss the caller builds on it) and returns the
//! vector-Jacobian products for the inflow controls, the buoyancy
//! coefficient and the initial state, from the library's hand-written
//! adjoint kernels.  No counterpart in the reference.

use crate::ffi;
use crate::fluid::{check, FluidSolver3D};

/// The final (or initial) state of a rollout: u, v, w, density,
/// temperature, each nx*ny*nz values in the solver's layout.
#[derive(Clone, Debug, Default)]
pub struct RolloutState {
    pub u: Vec<f64>,
    pub v: Vec<f64>,
    pub w: Vec<f64>,
    pub density: Vec<f64>,
    pub temperature: Vec<f64>,
}

impl RolloutState {
    fn from_flat(flat: &[f64], n: usize) -> Self {
        Self {
            u: flat[0..n].to_vec(),
            v: flat[n..2 * n].to_vec(),
            w: flat[2 * n..3 * n].to_vec(),
            density: flat[3 * n..4 * n].to_vec(),
            temperature: flat[4 * n..5 * n].to_vec(),
        }
    }

    fn to_flat(&self, n: usize) -> Vec<f64> {
        let mut flat = vec![0.0; 5 * n];
        for (q, f) in [&self.u, &self.v, &self.w, &self.density, &self.temperature]
            .into_iter()
            .enumerate()
        {
            if !f.is_empty() {
                assert_eq!(f.len(), n, "a state field has nx*ny*nz values");
                flat[q * n..(q + 1) * n].copy_from_slice(f);
            }
        }
        flat
    }
}

/// Gradients returned by [`SmokeTape::backward`].
#[derive(Clone, Debug, Default)]
pub struct RolloutGrads {
    pub controls: Vec<f64>,
    pub beta: f64,
    pub initial: RolloutState,
}

/// A rollout of `nsteps` steps of `dt` with density and temperature set
/// from per-cell controls on the inflow patch `[patch_lo, patch_hi)` (x
/// fastest) and buoyancy coefficient `beta`, from the solver's current
/// state, which forward and backward both leave unchanged.
pub struct SmokeTape<'a> {
    solver: &'a mut FluidSolver3D,
    desc: ffi::aether_rollout_desc,
    ncontrols: usize,
    recorded: bool,
}

impl<'a> SmokeTape<'a> {
    pub fn new(solver: &'a mut FluidSolver3D, nsteps: usize, dt: f64, patch_lo: [usize; 3],
               patch_hi: [usize; 3], checkpoint_every: usize) -> Self {
        let desc = ffi::aether_rollout_desc {
            nsteps: nsteps as i32,
            checkpoint_every: checkpoint_every as i32,
            dt,
            patch_lo: [patch_lo[0] as i32, patch_lo[1] as i32, patch_lo[2] as i32],
            patch_hi: [patch_hi[0] as i32, patch_hi[1] as i32, patch_hi[2] as i32],
            adjoint_cg_iters: solver.solver_iters as i32,
            reserved: 0,
            adjoint_cg_rel_tol: solver.params.cg_rel_tol,
        };
        let ncontrols = (0..3).map(|a| patch_hi[a].saturating_sub(patch_lo[a])).product();
        Self { solver, desc, ncontrols, recorded: false }
    }

    /// Projection solve of the adjoint (by default the forward's).
    pub fn adjoint_cg(mut self, max_iters: usize, rel_tol: f64) -> Self {
        self.desc.adjoint_cg_iters = max_iters as i32;
        self.desc.adjoint_cg_rel_tol = rel_tol;
        self
    }

    /// Runs the forward sweep and records the tape; returns the final state.
    pub fn forward(&mut self, controls: &[f64], beta: f64) -> RolloutState {
        assert_eq!(controls.len(), self.ncontrols, "one control per patch cell");
        let n = self.solver.nx * self.solver.ny * self.solver.nz;
        let mut flat = vec![0.0; 5 * n];
        // the knobs and any host-modified arrays reach the device first
        let _ = self.solver.compute_telemetry(0.0);
        check(unsafe {
            ffi::aether_fluid_rollout_forward(self.solver.dev, &self.desc, controls.as_ptr(), beta,
                                              flat.as_mut_ptr())
        });
        self.recorded = true;
        RolloutState::from_flat(&flat, n)
    }

    /// Consumes the tape: `lam_final` is the adjoint of the final state
    /// (empty fields count as zero).
    pub fn backward(&mut self, lam_final: &RolloutState) -> RolloutGrads {
        assert!(self.recorded, "backward() needs a forward() before it");
        self.recorded = false;
        let n = self.solver.nx * self.solver.ny * self.solver.nz;
        let lam = lam_final.to_flat(n);
        let mut gc = vec![0.0; self.ncontrols];
        let mut gb = 0.0;
        let mut gi = vec![0.0; 5 * n];
        check(unsafe {
            ffi::aether_fluid_rollout_backward(self.solver.dev, lam.as_ptr(), gc.as_mut_ptr(),
                                               &mut gb, gi.as_mut_ptr())
        });
        RolloutGrads { controls: gc, beta: gb, initial: RolloutState::from_flat(&gi, n) }
    }

    /// The fused form for loss = sum (density_T - target)^2: (loss,
    /// d/dcontrols, d/dbeta) in one call.
    pub fn loss_and_grad(&mut self, controls: &[f64], beta: f64, target: &[f64]) -> (f64, Vec<f64>, f64) {
        assert_eq!(controls.len(), self.ncontrols, "one control per patch cell");
        let _ = self.solver.compute_telemetry(0.0);
        let mut loss = 0.0;
        let mut gc = vec![0.0; self.ncontrols];
        let mut gb = 0.0;
        check(unsafe {
            ffi::aether_fluid_rollout(self.solver.dev, &self.desc, controls.as_ptr(), beta,
                                      target.as_ptr(), &mut loss, gc.as_mut_ptr(), &mut gb)
        });
        self.recorded = false;
        (loss, gc, gb)
    }
}
