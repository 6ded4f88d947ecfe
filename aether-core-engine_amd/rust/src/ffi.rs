//! `extern "C"` declarations of `include/aether_fluid.h`, one per entry
//! point, with the header's structs as `#[repr(C)]` types.  Every function
//! returns 0 or a negative `AETHER_E*` code; `aether_fluid_last_error` holds
//! the message.

#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_long, c_void};

/// `AETHER_FLUID_ABI_VERSION` of include/aether_fluid.h that the structs below follow.
pub const AETHER_FLUID_ABI_VERSION: c_int = 3;

pub const AETHER_OK: c_int = 0;
pub const AETHER_EINVAL: c_int = -1;
pub const AETHER_EHIP: c_int = -2;
pub const AETHER_ENOMEM: c_int = -3;
pub const AETHER_ENODEV: c_int = -4;
pub const AETHER_EINTERNAL: c_int = -5;

pub const AETHER_FP32: c_int = 0;
pub const AETHER_FP64: c_int = 1;
pub const AETHER_FP32_IR64: c_int = 2;

pub const AETHER_SOLVER_CG: i32 = 0;
pub const AETHER_SOLVER_RBGS: i32 = 1;

pub const AETHER_CG_AUTO: i32 = 0;
pub const AETHER_CG_SCALAR: i32 = 1;
pub const AETHER_CG_X4: i32 = 2;
pub const AETHER_CG_FUSED: i32 = 3;

pub const AETHER_SOURCE_JETS: i32 = 0;
pub const AETHER_SOURCE_SPHERE: i32 = 1;
pub const AETHER_SOURCE_NONE: i32 = 2;

pub const AETHER_FIELD_U: c_int = 0;
pub const AETHER_FIELD_V: c_int = 1;
pub const AETHER_FIELD_W: c_int = 2;
pub const AETHER_FIELD_DENSITY: c_int = 3;
pub const AETHER_FIELD_TEMPERATURE: c_int = 4;
pub const AETHER_FIELD_PRESSURE: c_int = 5;
pub const AETHER_FIELD_DIVERGENCE: c_int = 6;

#[repr(C)]
pub struct aether_fluid_s {
    _p: [u8; 0],
}
#[repr(C)]
pub struct aether_fluid_medium_s {
    _p: [u8; 0],
}
#[repr(C)]
pub struct aether_fluid_batch_s {
    _p: [u8; 0],
}
pub type aether_fluid_t = *mut aether_fluid_s;
pub type aether_fluid_medium_t = *mut aether_fluid_medium_s;
pub type aether_fluid_batch_t = *mut aether_fluid_batch_s;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_fluid_params {
    pub vorticity_eps: f64,
    pub buoyancy_beta: f64,
    pub gravity_alpha: f64,
    pub solver_kind: i32,
    pub solver_iters: i32,
    pub sor_omega: f64,
    pub cg_rel_tol: f64,
    pub dissipation_velocity: f64,
    pub dissipation_density: f64,
    pub dissipation_temperature: f64,
    pub source_kind: i32,
    pub cg_form: i32,
    pub source_center: [f64; 3],
    pub source_radius: f64,
    pub source_density: f64,
    pub source_temperature: f64,
    pub ir_inner_rel_tol: f64,
    pub ir_div_tol: f64,
    pub ir_max_outer: i32,
    pub batch_graph: i32,
    pub cg_z_cells: i32,
    pub reserved: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_fluid_telemetry {
    pub time_sec: f64,
    pub kinetic_energy: f64,
    pub enstrophy: f64,
    pub peak_vorticity: f64,
    pub total_mass: f64,
    pub center_of_mass: [f64; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_fluid_solve_stats {
    pub iterations: i32,
    pub converged: i32,
    pub b_norm2: f64,
    pub r_norm2_initial: f64,
    pub r_norm2_final: f64,
    pub outer_iterations: i32,
    pub reserved: i32,
    pub div_max_final: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_fluid_instance_desc {
    pub nx: i32,
    pub ny: i32,
    pub nz: i32,
    pub reserved: i32,
    pub h: f64,
    pub vorticity_eps: f64,
    pub buoyancy_beta: f64,
    pub gravity_alpha: f64,
    pub source_center: [f64; 3],
    pub source_radius: f64,
    pub source_density: f64,
    pub source_temperature: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_rollout_desc {
    pub nsteps: i32,
    pub checkpoint_every: i32,
    pub dt: f64,
    pub patch_lo: [i32; 3],
    pub patch_hi: [i32; 3],
    pub adjoint_cg_iters: i32,
    pub reserved: i32,
    pub adjoint_cg_rel_tol: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_medium_ray {
    pub origin: [f64; 3],
    pub dir: [f64; 3],
    pub t_min: f64,
    pub t_far: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct aether_medium_interaction {
    pub hit: i32,
    pub reserved: i32,
    pub t: f64,
    pub point: [f64; 3],
    pub albedo: [f64; 4],
    pub emission: [f64; 4],
}

extern "C" {
    // library and handle
    pub fn aether_fluid_abi_version() -> c_int;
    pub fn aether_fluid_last_error() -> *const c_char;
    pub fn aether_fluid_device_count(count: *mut c_int) -> c_int;
    pub fn aether_fluid_create(nx: c_int, ny: c_int, nz: c_int, h: f64, precision: c_int,
                               out: *mut aether_fluid_t) -> c_int;
    pub fn aether_fluid_destroy(s: aether_fluid_t) -> c_int;
    pub fn aether_fluid_dims(s: aether_fluid_t, nx: *mut c_int, ny: *mut c_int, nz: *mut c_int,
                             h: *mut f64, precision: *mut c_int) -> c_int;
    pub fn aether_fluid_default_params(out: *mut aether_fluid_params) -> c_int;
    pub fn aether_fluid_set_params(s: aether_fluid_t, p: *const aether_fluid_params) -> c_int;
    pub fn aether_fluid_get_params(s: aether_fluid_t, p: *mut aether_fluid_params) -> c_int;

    // FluidSolver3D methods (src/fluid.rs:88-672)
    pub fn aether_fluid_step(s: aether_fluid_t, dt: f64, time_sec: f64,
                             out: *mut aether_fluid_telemetry) -> c_int;
    pub fn aether_fluid_run(s: aether_fluid_t, dt: f64, times: *const f64, nsteps: c_int,
                            out: *mut aether_fluid_telemetry) -> c_int;
    pub fn aether_fluid_advect_maccormack_in_place(s: aether_fluid_t, field_kind: c_int, dt: f64,
                                                   dissipation: f64) -> c_int;
    pub fn aether_fluid_enforce_wall_velocity_bc(s: aether_fluid_t) -> c_int;
    pub fn aether_fluid_apply_buoyancy_and_vorticity_confinement(s: aether_fluid_t,
                                                                 dt: f64) -> c_int;
    pub fn aether_fluid_compute_divergence_norm(s: aether_fluid_t, rms: *mut f64) -> c_int;
    pub fn aether_fluid_project(s: aether_fluid_t) -> c_int;
    pub fn aether_fluid_solve_poisson_conjugate_gradient(s: aether_fluid_t,
                                                         max_iters: c_int) -> c_int;
    pub fn aether_fluid_solve_poisson_gauss_seidel(s: aether_fluid_t, sweeps: c_int) -> c_int;
    pub fn aether_fluid_inject_sources(s: aether_fluid_t, time_sec: f64, dt: f64) -> c_int;
    pub fn aether_fluid_compute_telemetry(s: aether_fluid_t, time_sec: f64,
                                          out: *mut aether_fluid_telemetry) -> c_int;
    pub fn aether_fluid_compute_telemetry_noalloc(precision: c_int, nx: c_int, ny: c_int,
                                                  nz: c_int, h: f64, u: *const c_void,
                                                  v: *const c_void, w: *const c_void,
                                                  density: *const c_void, time_sec: f64,
                                                  out_dev: *mut f64,
                                                  stream: *mut c_void) -> c_int;
    pub fn aether_fluid_sample_trilinear(s: aether_fluid_t, field: c_int, xyz: *const f64,
                                         npts: usize, out: *mut f64) -> c_int;
    pub fn aether_fluid_sample_trilinear_field(s: aether_fluid_t, field: *const f64, n: usize,
                                               xyz: *const f64, npts: usize,
                                               out: *mut f64) -> c_int;
    pub fn aether_fluid_divergence_max(s: aether_fluid_t, out: *mut f64) -> c_int;
    pub fn aether_fluid_last_solve_stats(s: aether_fluid_t,
                                         out: *mut aether_fluid_solve_stats) -> c_int;
    pub fn aether_fluid_last_run_iterations(s: aether_fluid_t, out: *mut i32, n: c_int,
                                            count: *mut c_int) -> c_int;
    pub fn aether_fluid_nonfinite(s: aether_fluid_t, flag: *mut c_int) -> c_int;

    // the pub Vec<f64> fields (src/fluid.rs:28-34)
    pub fn aether_fluid_upload(s: aether_fluid_t, field: c_int, src: *const f64,
                               n: usize) -> c_int;
    pub fn aether_fluid_download(s: aether_fluid_t, field: c_int, dst: *mut f64,
                                 n: usize) -> c_int;
    pub fn aether_fluid_max_density(s: aether_fluid_t, out: *mut f64) -> c_int;
    pub fn aether_fluid_autotune_cg(s: aether_fluid_t, iters: c_int, us_per_iter: *mut f64,
                                    form: *mut c_int, z_cells: *mut c_int) -> c_int;

    // FluidMedium (src/fluid.rs:687-910)
    pub fn aether_fluid_medium_create(s: aether_fluid_t, bounds_min: *const f64,
                                      bounds_max: *const f64, sigma_t_scale: f64, phase_g: f64,
                                      out: *mut aether_fluid_medium_t) -> c_int;
    pub fn aether_fluid_medium_destroy(m: aether_fluid_medium_t) -> c_int;
    pub fn aether_fluid_medium_info(m: aether_fluid_medium_t, max_density: *mut f64,
                                    majorant_sigma_t: *mut f64, phase_g: *mut f64) -> c_int;
    pub fn aether_fluid_medium_sample(m: aether_fluid_medium_t, xyz: *const f64, npts: usize,
                                      out: *mut f64) -> c_int;
    pub fn aether_fluid_medium_sample_interaction(m: aether_fluid_medium_t,
                                                  rays: *const aether_medium_ray, nrays: usize,
                                                  lambda: *const f64, rng_state: *mut u64,
                                                  out: *mut aether_medium_interaction) -> c_int;
    pub fn aether_fluid_medium_transmittance(m: aether_fluid_medium_t,
                                             rays: *const aether_medium_ray, nrays: usize,
                                             rng_state: *mut u64, out: *mut f64) -> c_int;
    pub fn aether_fluid_medium_phase_hg_pdf(m: aether_fluid_medium_t, wo: *const f64,
                                            wi: *const f64, n: usize, pdf: *mut f64) -> c_int;
    pub fn aether_fluid_medium_sample_phase_hg(m: aether_fluid_medium_t, wo: *const f64,
                                               n: usize, rng_state: *mut u64, wi: *mut f64,
                                               pdf: *mut f64) -> c_int;

    // src/rng.rs
    pub fn aether_rng_seed(seed: u64, state: *mut u64) -> c_int;
    pub fn aether_rng_fill_range(state: *mut u64, lo: f64, hi: f64, out: *mut f64,
                                 n: usize) -> c_int;

    // batched grids
    pub fn aether_fluid_batch_create(desc: *const aether_fluid_instance_desc, count: c_int,
                                     precision: c_int, shared: *const aether_fluid_params,
                                     out: *mut aether_fluid_batch_t) -> c_int;
    pub fn aether_fluid_batch_destroy(b: aether_fluid_batch_t) -> c_int;
    pub fn aether_fluid_batch_upload(b: aether_fluid_batch_t, inst: c_int, field: c_int,
                                     src: *const f64, n: usize) -> c_int;
    pub fn aether_fluid_batch_download(b: aether_fluid_batch_t, inst: c_int, field: c_int,
                                       dst: *mut f64, n: usize) -> c_int;
    pub fn aether_fluid_batch_run(b: aether_fluid_batch_t, dt: f64, times: *const f64,
                                  nsteps: c_int, out: *mut aether_fluid_telemetry) -> c_int;
    pub fn aether_fluid_batch_iterations(b: aether_fluid_batch_t, out: *mut i32, n: c_int,
                                         count: *mut c_int) -> c_int;
    pub fn aether_fluid_batch_autotune_cg(b: aether_fluid_batch_t, iters: c_int,
                                          us_per_iter: *mut f64, form: *mut c_int,
                                          z_cells: *mut c_int) -> c_int;

    // differentiable rollouts
    pub fn aether_fluid_rollout(s: aether_fluid_t, d: *const aether_rollout_desc,
                                controls: *const f64, beta: f64, target: *const f64,
                                loss: *mut f64, grad_controls: *mut f64,
                                grad_beta: *mut f64) -> c_int;
    pub fn aether_fluid_rollout_forward(s: aether_fluid_t, d: *const aether_rollout_desc,
                                        controls: *const f64, beta: f64,
                                        final_state: *mut f64) -> c_int;
    pub fn aether_fluid_rollout_backward(s: aether_fluid_t, lam_final: *const f64,
                                         grad_controls: *mut f64, grad_beta: *mut f64,
                                         grad_initial: *mut f64) -> c_int;

    // measurement
    pub fn aether_fluid_set_kernel_timing(s: aether_fluid_t, enable: c_int) -> c_int;
    pub fn aether_fluid_stream_copy(nbytes: usize, reps: c_int, avg_ms: *mut f64) -> c_int;
    pub fn aether_fluid_kernel_timing(s: aether_fluid_t, which: c_int, avg_ms: *mut f64,
                                      launches: *mut c_long) -> c_int;
}
