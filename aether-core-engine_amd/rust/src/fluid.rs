//! The reference's `FluidSolver3D` and `FluidMedium` (src/fluid.rs:6-910)
//! with their public interface kept, backed by the HIP library.
//!
//! Differences a caller sees, all at the type level:
//! * the public `Vec<f64>` fields (`u`, `v`, `w`, `density`, `temperature`,
//!   `pressure`, `divergence`, src/fluid.rs:28-34) become accessors: `u()`
//!   downloads the device field on first use, `u_mut()` marks it dirty so it
//!   is uploaded before the next solver call;
//! * `Vec3`, `Aabb`, `Ray`, `Wavelengths4` and the RNG state of the
//!   reference's math module are passed as plain arrays (`[f64; 3]`,
//!   `[f64; 4]`, `[u64; 4]`); inside the reference crate a `From` impl per
//!   type bridges them.
//!
//! Misuse panics, as the reference does (`assert!`, `fluid.rs:54`): every
//! non-zero status of the C ABI becomes a panic carrying the library's
//! message.

use std::cell::{Cell, Ref, RefCell, RefMut};
use std::ffi::CStr;

use crate::ffi;

/// src/fluid.rs:6-15
#[derive(Clone, Copy, Debug, Default)]
pub struct FluidTelemetryFrame {
    pub time_sec: f64,
    pub kinetic_energy: f64,
    pub enstrophy: f64,
    pub peak_vorticity: f64,
    pub total_mass: f64,
    /// Plume centre of mass normalised to [-1, 1]^3.
    pub center_of_mass: [f64; 3],
}

impl From<ffi::aether_fluid_telemetry> for FluidTelemetryFrame {
    fn from(t: ffi::aether_fluid_telemetry) -> Self {
        Self {
            time_sec: t.time_sec,
            kinetic_energy: t.kinetic_energy,
            enstrophy: t.enstrophy,
            peak_vorticity: t.peak_vorticity,
            total_mass: t.total_mass,
            center_of_mass: t.center_of_mass,
        }
    }
}

/// src/fluid.rs:17-21
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum PoissonSolverKind {
    ConjugateGradient,
    RedBlackGaussSeidel,
}

/// Device precision of a solver.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Precision {
    /// FP64 everywhere: the reference's arithmetic (the drop-in default).
    Fp64,
    /// FP32 fields and stencils, FP64 reductions: the throughput setting.
    Fp32,
    /// FP32 fields with an FP64-refined pressure solve.
    Fp32Refined,
}

pub(crate) fn check(rc: i32) {
    if rc != ffi::AETHER_OK {
        let msg = unsafe { CStr::from_ptr(ffi::aether_fluid_last_error()) };
        panic!("aether_fluid: {} (status {})", msg.to_string_lossy(), rc);
    }
}

/// The seven public arrays of the reference's solver, in the library's
/// field-id order.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Field {
    U = 0,
    V = 1,
    W = 2,
    Density = 3,
    Temperature = 4,
    Pressure = 5,
    Divergence = 6,
}

/// src/fluid.rs:23-50.  The knob fields are the reference's; the arrays
/// live on the device.
pub struct FluidSolver3D {
    pub nx: usize,
    pub ny: usize,
    pub nz: usize,
    pub h: f64,
    pub vorticity_eps: f64,
    pub buoyancy_beta: f64,
    pub gravity_alpha: f64,
    pub solver_kind: PoissonSolverKind,
    pub solver_iters: usize,
    /// Extensions with the reference's values as defaults (sor_omega 1.65,
    /// no relative tolerance, dissipation 0.996 / 0.996 / 0.985, the jets).
    pub params: ffi::aether_fluid_params,
    pub(crate) dev: ffi::aether_fluid_t,
    // host copies of the seven arrays, one cell each so that reading one
    // field while another is downloaded is allowed
    mirrors: [RefCell<Vec<f64>>; 7],
    valid: Cell<[bool; 7]>,
    dirty: Cell<[bool; 7]>,
}

impl FluidSolver3D {
    /// src/fluid.rs:53-85 (FP64, the reference's arithmetic).
    pub fn new(nx: usize, ny: usize, nz: usize) -> Self {
        Self::with_precision(nx, ny, nz, Precision::Fp64)
    }

    pub fn with_precision(nx: usize, ny: usize, nz: usize, precision: Precision) -> Self {
        assert!(nx >= 6 && ny >= 6 && nz >= 6); // src/fluid.rs:54
        let prec = match precision {
            Precision::Fp64 => ffi::AETHER_FP64,
            Precision::Fp32 => ffi::AETHER_FP32,
            Precision::Fp32Refined => ffi::AETHER_FP32_IR64,
        };
        let abi = unsafe { ffi::aether_fluid_abi_version() };
        assert_eq!(abi, ffi::AETHER_FLUID_ABI_VERSION, "libaether_fluid_hip.so has another ABI version");
        let mut dev: ffi::aether_fluid_t = std::ptr::null_mut();
        check(unsafe { ffi::aether_fluid_create(nx as i32, ny as i32, nz as i32, 0.0, prec, &mut dev) });
        let mut params = ffi::aether_fluid_params::default();
        check(unsafe { ffi::aether_fluid_default_params(&mut params) });
        let (mut ix, mut iy, mut iz, mut h, mut p) = (0, 0, 0, 0.0, 0);
        check(unsafe { ffi::aether_fluid_dims(dev, &mut ix, &mut iy, &mut iz, &mut h, &mut p) });
        Self {
            nx,
            ny,
            nz,
            h,
            vorticity_eps: params.vorticity_eps, // 5.5, src/fluid.rs:79-83
            buoyancy_beta: params.buoyancy_beta, // 6.8
            gravity_alpha: params.gravity_alpha, // 0.45
            solver_kind: PoissonSolverKind::ConjugateGradient,
            solver_iters: params.solver_iters as usize, // 32
            params,
            dev,
            mirrors: Default::default(),
            valid: Cell::new([false; 7]),
            dirty: Cell::new([false; 7]),
        }
    }

    /// src/fluid.rs:88-90
    #[inline]
    pub fn idx(&self, i: usize, j: usize, k: usize) -> usize {
        (k * self.ny + j) * self.nx + i
    }

    fn cells(&self) -> usize {
        self.nx * self.ny * self.nz
    }

    /// Pushes the knob fields and any host-modified arrays to the device.
    fn sync_in(&self) {
        let mut p = self.params;
        p.vorticity_eps = self.vorticity_eps;
        p.buoyancy_beta = self.buoyancy_beta;
        p.gravity_alpha = self.gravity_alpha;
        p.solver_kind = match self.solver_kind {
            PoissonSolverKind::ConjugateGradient => ffi::AETHER_SOLVER_CG,
            PoissonSolverKind::RedBlackGaussSeidel => ffi::AETHER_SOLVER_RBGS,
        };
        p.solver_iters = self.solver_iters as i32;
        check(unsafe { ffi::aether_fluid_set_params(self.dev, &p) });
        let n = self.cells();
        let mut dirty = self.dirty.get();
        for (id, d) in dirty.iter_mut().enumerate() {
            if *d {
                let data = self.mirrors[id].borrow();
                check(unsafe { ffi::aether_fluid_upload(self.dev, id as i32, data.as_ptr(), n) });
                *d = false;
            }
        }
        self.dirty.set(dirty);
    }

    /// After a device call that may change the arrays.
    fn invalidate(&self) {
        self.valid.set([false; 7]);
    }

    fn mirror(&self, f: Field) -> Ref<'_, [f64]> {
        let id = f as usize;
        let mut valid = self.valid.get();
        if !valid[id] {
            let mut data = self.mirrors[id].borrow_mut();
            data.resize(self.cells(), 0.0);
            check(unsafe { ffi::aether_fluid_download(self.dev, f as i32, data.as_mut_ptr(), data.len()) });
            valid[id] = true;
            self.valid.set(valid);
        }
        Ref::map(self.mirrors[id].borrow(), |d| d.as_slice())
    }

    fn mirror_mut(&mut self, f: Field) -> RefMut<'_, [f64]> {
        let id = f as usize;
        drop(self.mirror(f));
        let mut dirty = self.dirty.get();
        dirty[id] = true;
        self.dirty.set(dirty);
        RefMut::map(self.mirrors[id].borrow_mut(), |d| d.as_mut_slice())
    }

    /// The reference's `pub u: Vec<f64>` (and the other arrays), read.
    pub fn field(&self, f: Field) -> Ref<'_, [f64]> {
        self.mirror(f)
    }
    /// ... and written (uploaded before the next solver call).
    pub fn field_mut(&mut self, f: Field) -> RefMut<'_, [f64]> {
        self.mirror_mut(f)
    }
    pub fn u(&self) -> Ref<'_, [f64]> { self.mirror(Field::U) }
    pub fn v(&self) -> Ref<'_, [f64]> { self.mirror(Field::V) }
    pub fn w(&self) -> Ref<'_, [f64]> { self.mirror(Field::W) }
    pub fn density(&self) -> Ref<'_, [f64]> { self.mirror(Field::Density) }
    pub fn temperature(&self) -> Ref<'_, [f64]> { self.mirror(Field::Temperature) }
    pub fn pressure(&self) -> Ref<'_, [f64]> { self.mirror(Field::Pressure) }
    pub fn divergence(&self) -> Ref<'_, [f64]> { self.mirror(Field::Divergence) }
    pub fn u_mut(&mut self) -> RefMut<'_, [f64]> { self.mirror_mut(Field::U) }
    pub fn v_mut(&mut self) -> RefMut<'_, [f64]> { self.mirror_mut(Field::V) }
    pub fn w_mut(&mut self) -> RefMut<'_, [f64]> { self.mirror_mut(Field::W) }
    pub fn density_mut(&mut self) -> RefMut<'_, [f64]> { self.mirror_mut(Field::Density) }
    pub fn temperature_mut(&mut self) -> RefMut<'_, [f64]> { self.mirror_mut(Field::Temperature) }

    /// src/fluid.rs:94-128, on any field of this grid's layout.
    pub fn sample_trilinear(&self, field: &[f64], x: f64, y: f64, z: f64) -> f64 {
        let xyz = [x, y, z];
        let mut out = 0.0;
        check(unsafe {
            ffi::aether_fluid_sample_trilinear_field(self.dev, field.as_ptr(), field.len(),
                                                     xyz.as_ptr(), 1, &mut out)
        });
        out
    }

    /// src/fluid.rs:170-234
    pub fn advect_maccormack_in_place(&mut self, field_kind: u8, dt: f64, dissipation: f64) {
        self.sync_in();
        check(unsafe {
            ffi::aether_fluid_advect_maccormack_in_place(self.dev, field_kind as i32, dt, dissipation)
        });
        self.invalidate();
    }

    /// src/fluid.rs:236-253
    pub fn enforce_wall_velocity_bc(&mut self) {
        self.sync_in();
        check(unsafe { ffi::aether_fluid_enforce_wall_velocity_bc(self.dev) });
        self.invalidate();
    }

    /// src/fluid.rs:256-330
    pub fn apply_buoyancy_and_vorticity_confinement(&mut self, dt: f64) {
        self.sync_in();
        check(unsafe { ffi::aether_fluid_apply_buoyancy_and_vorticity_confinement(self.dev, dt) });
        self.invalidate();
    }

    /// src/fluid.rs:333-353 (also stores `divergence`)
    pub fn compute_divergence_norm(&mut self) -> f64 {
        self.sync_in();
        let mut rms = 0.0;
        check(unsafe { ffi::aether_fluid_compute_divergence_norm(self.dev, &mut rms) });
        self.invalidate();
        rms
    }

    /// src/fluid.rs:397-442
    pub fn solve_poisson_gauss_seidel(&mut self, iterations: usize) {
        self.sync_in();
        check(unsafe { ffi::aether_fluid_solve_poisson_gauss_seidel(self.dev, iterations as i32) });
        self.invalidate();
    }

    /// src/fluid.rs:445-493
    pub fn solve_poisson_conjugate_gradient(&mut self, max_iters: usize) {
        self.sync_in();
        check(unsafe { ffi::aether_fluid_solve_poisson_conjugate_gradient(self.dev, max_iters as i32) });
        self.invalidate();
    }

    /// src/fluid.rs:496-529
    pub fn project(&mut self) {
        self.sync_in();
        check(unsafe { ffi::aether_fluid_project(self.dev) });
        self.invalidate();
    }

    /// src/fluid.rs:532-595
    pub fn inject_cyberpunk_sources(&mut self, time_sec: f64, dt: f64) {
        self.sync_in();
        check(unsafe { ffi::aether_fluid_inject_sources(self.dev, time_sec, dt) });
        self.invalidate();
    }

    /// src/fluid.rs:598-611
    pub fn step(&mut self, sim_dt: f64, timeline_sec: f64) -> FluidTelemetryFrame {
        self.sync_in();
        let mut tel = ffi::aether_fluid_telemetry::default();
        check(unsafe { ffi::aether_fluid_step(self.dev, sim_dt, timeline_sec, &mut tel) });
        self.invalidate();
        tel.into()
    }

    /// The src/main.rs:89-94 loop in one call (one host synchronisation):
    /// step i at `times[i]`.
    pub fn run(&mut self, sim_dt: f64, times: &[f64]) -> Vec<FluidTelemetryFrame> {
        self.sync_in();
        let mut tel = vec![ffi::aether_fluid_telemetry::default(); times.len()];
        check(unsafe {
            ffi::aether_fluid_run(self.dev, sim_dt, times.as_ptr(), times.len() as i32, tel.as_mut_ptr())
        });
        self.invalidate();
        tel.into_iter().map(Into::into).collect()
    }

    /// src/fluid.rs:613-672
    pub fn compute_telemetry(&self, timeline_sec: f64) -> FluidTelemetryFrame {
        self.sync_in();
        let mut tel = ffi::aether_fluid_telemetry::default();
        check(unsafe { ffi::aether_fluid_compute_telemetry(self.dev, timeline_sec, &mut tel) });
        tel.into()
    }

    /// Whether a telemetry reduction since the last call met a NaN or an
    /// infinity (no counterpart in the reference).
    pub fn nonfinite(&self) -> bool {
        let mut f = 0;
        check(unsafe { ffi::aether_fluid_nonfinite(self.dev, &mut f) });
        f != 0
    }

    /// CG statistics of the last pressure solve (no counterpart in the reference).
    pub fn last_solve_stats(&self) -> ffi::aether_fluid_solve_stats {
        let mut s = ffi::aether_fluid_solve_stats::default();
        check(unsafe { ffi::aether_fluid_last_solve_stats(self.dev, &mut s) });
        s
    }
}

impl Drop for FluidSolver3D {
    fn drop(&mut self) {
        unsafe {
            ffi::aether_fluid_destroy(self.dev);
        }
    }
}

/// src/fluid.rs:679-685
#[derive(Clone, Copy, Debug, Default)]
pub struct MediumInteraction {
    pub t: f64,
    pub point: [f64; 3],
    pub albedo: [f64; 4],
    pub emission: [f64; 4],
}

/// A ray as the reference's `Ray::new` leaves it (src/math.rs:763-773): `dir`
/// normalised.
#[derive(Clone, Copy, Debug)]
pub struct Ray {
    pub origin: [f64; 3],
    pub dir: [f64; 3],
    pub t_min: f64,
}

/// src/fluid.rs:687-910: a device snapshot of the solver's density and
/// temperature, sampled by the path tracer.
pub struct FluidMedium {
    pub bounds_min: [f64; 3],
    pub bounds_max: [f64; 3],
    pub nx: usize,
    pub ny: usize,
    pub nz: usize,
    pub max_density: f64,
    pub sigma_t_scale: f64,
    pub majorant_sigma_t: f64,
    pub phase_g: f64,
    dev: ffi::aether_fluid_medium_t,
}

impl FluidMedium {
    /// src/fluid.rs:703-735
    pub fn from_solver(solver: &FluidSolver3D, bounds_min: [f64; 3], bounds_max: [f64; 3],
                       sigma_t_scale: f64, phase_g: f64) -> Self {
        solver.sync_in();
        let mut dev: ffi::aether_fluid_medium_t = std::ptr::null_mut();
        check(unsafe {
            ffi::aether_fluid_medium_create(solver.dev, bounds_min.as_ptr(), bounds_max.as_ptr(),
                                            sigma_t_scale, phase_g, &mut dev)
        });
        let (mut md, mut mj, mut g) = (0.0, 0.0, 0.0);
        check(unsafe { ffi::aether_fluid_medium_info(dev, &mut md, &mut mj, &mut g) });
        Self {
            bounds_min,
            bounds_max,
            nx: solver.nx,
            ny: solver.ny,
            nz: solver.nz,
            max_density: md,
            sigma_t_scale,
            majorant_sigma_t: mj,
            phase_g: g,
            dev,
        }
    }

    /// src/fluid.rs:739-788
    pub fn sample_density_and_temp(&self, p: [f64; 3]) -> (f64, f64) {
        let mut out = [0.0; 2];
        check(unsafe { ffi::aether_fluid_medium_sample(self.dev, p.as_ptr(), 1, out.as_mut_ptr()) });
        (out[0], out[1])
    }

    fn ffi_ray(ray: &Ray, t_far: f64) -> ffi::aether_medium_ray {
        ffi::aether_medium_ray { origin: ray.origin, dir: ray.dir, t_min: ray.t_min, t_far }
    }

    /// src/fluid.rs:793-838 (`rng`: the ray's Xoshiro256++ state, advanced).
    pub fn sample_interaction(&self, ray: &Ray, surface_t: f64, lambda: &[f64; 4],
                              rng: &mut [u64; 4]) -> Option<MediumInteraction> {
        let r = Self::ffi_ray(ray, surface_t);
        let mut out = ffi::aether_medium_interaction::default();
        check(unsafe {
            ffi::aether_fluid_medium_sample_interaction(self.dev, &r, 1, lambda.as_ptr(),
                                                        rng.as_mut_ptr(), &mut out)
        });
        (out.hit != 0).then(|| MediumInteraction {
            t: out.t,
            point: out.point,
            albedo: out.albedo,
            emission: out.emission,
        })
    }

    /// The same for a batch of rays (one device launch): the renderer's
    /// rays of one bounce; rng holds one state per ray.
    pub fn sample_interactions(&self, rays: &[ffi::aether_medium_ray], lambda: &[[f64; 4]],
                               rng: &mut [[u64; 4]]) -> Vec<Option<MediumInteraction>> {
        assert!(lambda.len() == rays.len() && rng.len() == rays.len());
        let mut out = vec![ffi::aether_medium_interaction::default(); rays.len()];
        check(unsafe {
            ffi::aether_fluid_medium_sample_interaction(self.dev, rays.as_ptr(), rays.len(),
                                                        lambda.as_ptr() as *const f64,
                                                        rng.as_mut_ptr() as *mut u64,
                                                        out.as_mut_ptr())
        });
        out.iter()
            .map(|o| (o.hit != 0).then(|| MediumInteraction {
                t: o.t,
                point: o.point,
                albedo: o.albedo,
                emission: o.emission,
            }))
            .collect()
    }

    /// src/fluid.rs:841-874
    pub fn transmittance(&self, ray: &Ray, max_dist: f64, rng: &mut [u64; 4]) -> f64 {
        let r = Self::ffi_ray(ray, max_dist);
        let mut tr = 1.0;
        check(unsafe {
            ffi::aether_fluid_medium_transmittance(self.dev, &r, 1, rng.as_mut_ptr(), &mut tr)
        });
        tr
    }

    /// src/fluid.rs:878-888
    pub fn phase_hg_pdf(&self, wo: [f64; 3], wi: [f64; 3]) -> f64 {
        let mut pdf = 0.0;
        check(unsafe { ffi::aether_fluid_medium_phase_hg_pdf(self.dev, wo.as_ptr(), wi.as_ptr(), 1, &mut pdf) });
        pdf
    }

    /// src/fluid.rs:892-910
    pub fn sample_phase_hg(&self, wo: [f64; 3], rng: &mut [u64; 4]) -> ([f64; 3], f64) {
        let mut wi = [0.0; 3];
        let mut pdf = 0.0;
        check(unsafe {
            ffi::aether_fluid_medium_sample_phase_hg(self.dev, wo.as_ptr(), 1, rng.as_mut_ptr(),
                                                     wi.as_mut_ptr(), &mut pdf)
        });
        (wi, pdf)
    }
}

impl Drop for FluidMedium {
    fn drop(&mut self) {
        unsafe {
            ffi::aether_fluid_medium_destroy(self.dev);
        }
    }
}
