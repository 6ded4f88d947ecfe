/*
 * aether_fluid.h -- C ABI of the MI355X (gfx950) smoke solver.
 *
 * Drop-in boundary for the reference's `pub mod fluid` (src/lib.rs:4):
 * every entry point below replaces one public item of `FluidSolver3D`
 * (src/fluid.rs:23-673), cited beside it.  Plain C types only: opaque handle,
 * ints, doubles, host pointers and sizes.  Fields cross the boundary as f64
 * host arrays in the reference's layout, idx = (k*ny + j)*nx + i
 * (src/fluid.rs:88-90); the library keeps them on the device in FP32 or FP64.
 *
 * Every function returns 0 on success or a negative AETHER_E* code; the text
 * of the last error on the calling thread is returned by
 * aether_fluid_last_error().  The library never aborts the process.
 *
 * A handle is not thread-safe (the reference's methods take &mut self); two
 * handles may be driven concurrently from different threads.  Every call is
 * synchronous at return, like the reference's blocking methods.
 */
#ifndef AETHER_FLUID_H
#define AETHER_FLUID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AETHER_FLUID_ABI_VERSION 3

/* status codes */
#define AETHER_OK 0
#define AETHER_EINVAL (-1)   /* bad argument: dims < 6, null handle, bad field id, size mismatch */
#define AETHER_EHIP (-2)     /* HIP runtime error */
#define AETHER_ENOMEM (-3)   /* device allocation failed */
#define AETHER_ENODEV (-4)   /* no usable gfx950 device */
#define AETHER_EINTERNAL (-5) /* unexpected internal failure (message in aether_fluid_last_error) */

/* precision of the device state */
#define AETHER_FP32 0        /* FP32 fields and stencils, FP64 CG dot products */
#define AETHER_FP64 1        /* FP64 everywhere (parity instantiation) */
#define AETHER_FP32_IR64 2   /* FP32 fields; FP64 pressure/rhs with FP32 inner CG inside FP64 iterative refinement */

/* Poisson solver, src/fluid.rs:17-21 PoissonSolverKind */
#define AETHER_SOLVER_CG 0
#define AETHER_SOLVER_RBGS 1

/* CG iteration layout (all follow src/fluid.rs:463-487; in FP64 every layout
 * keeps the reference's per-cell evaluation order, and the layouts differ only
 * in the order their dot products are summed) */
#define AETHER_CG_AUTO 0     /* FP32: nx % 4 != 0: AETHER_CG_SCALAR; else AETHER_CG_FUSED up to
                                3.2 M cells (cache-resident), AETHER_CG_X4 above.  FP64:
                                AETHER_CG_X4 when nx % 2 == 0, else AETHER_CG_SCALAR.  Or
                                the aether_fluid_autotune_cg choice */
#define AETHER_CG_SCALAR 1   /* two launches per iteration, one cell per thread (also keeps
                                the red-black SOR passes at one cell per thread) */
#define AETHER_CG_X4 2       /* two launches per iteration, one 16-byte lane of x-cells per
                                thread: four in FP32 (nx % 4 == 0), two in FP64
                                (nx % 2 == 0) */
#define AETHER_CG_FUSED 3    /* one launch per iteration, single-reduction form (FP32; FP64
                                runs AETHER_CG_SCALAR);
                                batches (every nx % 4 == 0): solved group by group, each
                                group at most 6.4 M cells */

/* source kinds */
#define AETHER_SOURCE_JETS 0   /* src/fluid.rs:532-595 inject_cyberpunk_sources */
#define AETHER_SOURCE_SPHERE 1 /* sphere (cell-index coords, d^2<r^2, interior cells) that sets rho and T */
#define AETHER_SOURCE_NONE 2

/* field ids: the reference's pub arrays (src/fluid.rs:28-34) */
#define AETHER_FIELD_U 0
#define AETHER_FIELD_V 1
#define AETHER_FIELD_W 2
#define AETHER_FIELD_DENSITY 3
#define AETHER_FIELD_TEMPERATURE 4
#define AETHER_FIELD_PRESSURE 5
#define AETHER_FIELD_DIVERGENCE 6

typedef struct aether_fluid_s *aether_fluid_t;

/* Solver knobs.  The first five are the reference's pub fields
 * (src/fluid.rs:45-49, defaults src/fluid.rs:79-83); the rest are constants
 * the reference hard-codes, exposed with the reference's values as defaults. */
typedef struct {
    double vorticity_eps;           /* 5.5 */
    double buoyancy_beta;           /* 6.8 */
    double gravity_alpha;           /* 0.45 */
    int32_t solver_kind;            /* AETHER_SOLVER_CG */
    int32_t solver_iters;           /* 32: CG max iterations / RBGS sweeps */
    double sor_omega;               /* 1.65 (src/fluid.rs:398) */
    double cg_rel_tol;              /* 0: reference absolute guards only (src/fluid.rs:462,469,479);
                                       >0: also stop when ||r||2 <= tol*||b||2 */
    double dissipation_velocity;    /* 0.996 (src/fluid.rs:599-601) */
    double dissipation_density;     /* 0.996 (src/fluid.rs:607) */
    double dissipation_temperature; /* 0.985 (src/fluid.rs:608) */
    int32_t source_kind;            /* AETHER_SOURCE_JETS */
    int32_t cg_form;                /* AETHER_CG_AUTO */
    double source_center[3];        /* sphere source centre, cell-index coordinates */
    double source_radius;           /* sphere radius in cells */
    double source_density;          /* value set inside the sphere */
    double source_temperature;      /* value set inside the sphere */
    double ir_inner_rel_tol;        /* AETHER_FP32_IR64: inner FP32 CG tolerance (1e-3) */
    double ir_div_tol;              /* AETHER_FP32_IR64: stop when max|div u| <= this (1e-7) */
    int32_t ir_max_outer;           /* AETHER_FP32_IR64: max refinement sweeps (30) */
    int32_t batch_graph;            /* batches only: 1 = capture each step into a HIP graph
                                       and replay it (all solver_iters CG launches are in
                                       the graph; converged instances return at once) */
    int32_t cg_z_cells;             /* z cells per thread of the CG tiling: 0 = the
                                       layout's default or the autotuner's choice; the
                                       16-byte-lane two-launch layout takes 2 or 4, the
                                       one-cell-per-thread layout 4, 8 or 16 (batches: 8),
                                       the batches' one-launch layout 4 */
    int32_t reserved;
} aether_fluid_params;

/* src/fluid.rs:6-15 FluidTelemetryFrame (center_of_mass flattened) */
typedef struct {
    double time_sec;
    double kinetic_energy;
    double enstrophy;
    double peak_vorticity;
    double total_mass;
    double center_of_mass[3];
} aether_fluid_telemetry;

/* Statistics of the last pressure solve (no counterpart in the reference,
 * whose CG returns nothing: src/fluid.rs:445). */
typedef struct {
    int32_t iterations;      /* CG iterations that updated p (RBGS: sweeps) */
    int32_t converged;       /* 1 if a stopping test fired before max iterations */
    double b_norm2;          /* ||b||_2^2 of the right-hand side */
    double r_norm2_initial;  /* ||b - A p0||_2^2 */
    double r_norm2_final;    /* recursive residual ||r||_2^2 at exit */
    int32_t outer_iterations;/* AETHER_FP32_IR64 refinement sweeps (else 0) */
    int32_t reserved;
    double div_max_final;    /* AETHER_FP32_IR64: max|div u| reached (else 0) */
} aether_fluid_solve_stats;

/* Library / device information. */
int aether_fluid_abi_version(void);
const char *aether_fluid_last_error(void);
int aether_fluid_device_count(int *count);

/* src/fluid.rs:53-85 FluidSolver3D::new.  h <= 0 selects the reference rule
 * h = 1/max(nx,ny,nz) (src/fluid.rs:56); nx,ny,nz >= 6 (src/fluid.rs:54). */
int aether_fluid_create(int nx, int ny, int nz, double h, int precision, aether_fluid_t *out);
int aether_fluid_destroy(aether_fluid_t s);
int aether_fluid_dims(aether_fluid_t s, int *nx, int *ny, int *nz, double *h, int *precision);

/* Reference defaults for every knob (src/fluid.rs:79-83, 398, 599-608). */
int aether_fluid_default_params(aether_fluid_params *out);
/* Writes of the pub fields vorticity_eps, buoyancy_beta, gravity_alpha,
 * solver_kind, solver_iters (src/fluid.rs:45-49). */
int aether_fluid_set_params(aether_fluid_t s, const aether_fluid_params *p);
int aether_fluid_get_params(aether_fluid_t s, aether_fluid_params *p);

/* src/fluid.rs:598-611 FluidSolver3D::step.  `out` may be NULL. */
int aether_fluid_step(aether_fluid_t s, double dt, double time_sec, aether_fluid_telemetry *out);
/* Batched driver of the src/main.rs:89-94 loop: runs `nsteps` steps, step i
 * at time times[i], writing one telemetry frame per step into `out` (may be
 * NULL), with a single host synchronisation at the end. */
int aether_fluid_run(aether_fluid_t s, double dt, const double *times, int nsteps,
                     aether_fluid_telemetry *out);

/* Per-stage methods (the calls made by src/verify.rs:231-238 and the
 * fluid.rs tests). */
int aether_fluid_advect_maccormack_in_place(aether_fluid_t s, int field_kind, double dt, double dissipation); /* fluid.rs:170 */
int aether_fluid_enforce_wall_velocity_bc(aether_fluid_t s);                            /* fluid.rs:236 */
int aether_fluid_apply_buoyancy_and_vorticity_confinement(aether_fluid_t s, double dt);                     /* fluid.rs:256 */
int aether_fluid_compute_divergence_norm(aether_fluid_t s, double *rms);                        /* fluid.rs:333 */
int aether_fluid_project(aether_fluid_t s);                                             /* fluid.rs:496 */
int aether_fluid_solve_poisson_conjugate_gradient(aether_fluid_t s, int max_iters);                             /* fluid.rs:445 */
int aether_fluid_solve_poisson_gauss_seidel(aether_fluid_t s, int sweeps);                              /* fluid.rs:397 */
int aether_fluid_inject_sources(aether_fluid_t s, double time_sec, double dt);          /* fluid.rs:532 */
int aether_fluid_compute_telemetry(aether_fluid_t s, double time_sec, aether_fluid_telemetry *out); /* fluid.rs:613 */
/* No counterpart in the reference: a variant of aether_fluid_compute_telemetry
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It reduces
 * the telemetry (fluid.rs:613-672) of fields the caller holds on the device:
 * u, v, w, density are nx*ny*nz arrays of float (AETHER_FP32) or double
 * (AETHER_FP64) in the reference layout; h <= 0 selects 1/max(nx,ny,nz).
 * out_dev: 8 doubles on the device {time, ke, enstrophy, peak |w|, mass,
 * com x, y, z}; stream: a hipStream_t (NULL: the null stream).  It enqueues
 * two kernels and returns without synchronising; it calls no allocation
 * routine.  The partial sums go to a library-owned device array of 224 KiB
 * (4096 slots) that each GPU holds until the process ends: grids needing
 * more slots (more than 4096 columns of ceil(nx/64) x ceil(ny/4)) are
 * rejected with AETHER_EINVAL.  Calls that overlap on one GPU (different
 * streams or threads) share that array and race.  The frame equals
 * aether_fluid_compute_telemetry's bit for bit. */
int aether_fluid_compute_telemetry_noalloc(int precision, int nx, int ny, int nz, double h,
                                           const void *u, const void *v, const void *w,
                                           const void *density, double time_sec, double *out_dev,
                                           void *stream);
/* fluid.rs:94 sample_trilinear, for npts points (x,y,z interleaved), of one
 * of the solver's own fields (an AETHER_FIELD_* id). */
int aether_fluid_sample_trilinear(aether_fluid_t s, int field, const double *xyz, size_t npts,
                                  double *out);
/* fluid.rs:94 sample_trilinear(&self, field: &[f64], x, y, z) as the
 * reference declares it: any caller-supplied f64 field of n = nx*ny*nz
 * values in this grid's layout, evaluated in FP64 with the grid's clamp box,
 * at npts points (x,y,z interleaved). */
int aether_fluid_sample_trilinear_field(aether_fluid_t s, const double *field, size_t n,
                                        const double *xyz, size_t npts, double *out);
/* Extension: max |div u| over interior cells (same stencil as fluid.rs:341-344). */
int aether_fluid_divergence_max(aether_fluid_t s, double *out);
int aether_fluid_last_solve_stats(aether_fluid_t s, aether_fluid_solve_stats *out);
/* CG iterations of each step of the last aether_fluid_run / aether_fluid_step
 * (RBGS: sweeps); copies min(n, steps) values, *count = steps. */
int aether_fluid_last_run_iterations(aether_fluid_t s, int32_t *out, int n, int *count);

/* No counterpart in the reference, whose solver lets a NaN or infinity
 * propagate silently (src/fluid.rs:94-128 clamps and floors a NaN trace to
 * index 0).  Every telemetry reduction (step, run, compute_telemetry) raises
 * a device flag when a cell's velocity, |curl u|^2 or density is not finite;
 * *flag receives it (1 or 0) and it is cleared.  The fields themselves are
 * left as the reference would leave them. */
int aether_fluid_nonfinite(aether_fluid_t s, int *flag);

/* Field access (the reference's pub Vec<f64> arrays, src/fluid.rs:28-34).
 * n must equal nx*ny*nz. */
int aether_fluid_upload(aether_fluid_t s, int field, const double *src, size_t n);
int aether_fluid_download(aether_fluid_t s, int field, double *dst, size_t n);

/* src/fluid.rs:703-708 FluidMedium::from_solver: max density over all cells
 * (reduced on the device) and the f64 density/temperature snapshots. */
int aether_fluid_max_density(aether_fluid_t s, double *out);

/* Extension (no counterpart in the reference): CG layout autotuner.  Times
 * `iters` CG iterations of each candidate on a synthetic right-hand side of
 * this grid (pressure and divergence are saved and restored; the last-solve
 * statistics then describe the tuning solves) and makes the fastest the
 * AETHER_CG_AUTO choice of this handle.  us_per_iter (6 entries, may be
 * NULL): the 16-byte-lane layout with 2 and 4 z cells per thread, one cell
 * per thread with 4, 8 and 16, the one-launch form; -1 where a candidate does
 * not apply (FP64 has no one-launch form).  *form / *z_cells: the choice. */
int aether_fluid_autotune_cg(aether_fluid_t s, int iters, double *us_per_iter, int *form,
                             int *z_cells);

/* src/fluid.rs:687-788 FluidMedium, the renderer-side view of the solver.
 * aether_fluid_medium_create is FluidMedium::from_solver (fluid.rs:703-735):
 * it snapshots density and temperature on the device (later solver steps do
 * not change the medium), reduces max_density there, and sets
 * majorant_sigma_t = max(max_density * sigma_t_scale, 1e-4) and
 * phase_g = clamp(phase_g, -0.9, 0.9).  The spectral curves are renderer
 * constants and stay with the renderer. */
typedef struct aether_fluid_medium_s *aether_fluid_medium_t;
int aether_fluid_medium_create(aether_fluid_t s, const double *bounds_min,
                               const double *bounds_max, double sigma_t_scale, double phase_g,
                               aether_fluid_medium_t *out);
int aether_fluid_medium_destroy(aether_fluid_medium_t m);
int aether_fluid_medium_info(aether_fluid_medium_t m, double *max_density,
                             double *majorant_sigma_t, double *phase_g);
/* fluid.rs:739-788 sample_density_and_temp at npts points (x,y,z
 * interleaved), evaluated in FP64 on the device: out receives npts
 * (density, temperature) pairs, (0, 0) outside the bounds. */
int aether_fluid_medium_sample(aether_fluid_medium_t m, const double *xyz, size_t npts,
                               double *out);

/* FluidMedium's volumetric sampling for the path tracer (src/fluid.rs:790-910;
 * called per ray by src/renderer.rs:231-306), on the device in FP64 for a
 * batch of rays.  Every ray has its own Xoshiro256++ state (4 words, the
 * reference's Rng, src/rng.rs:14-77) in rng_state[4*i .. 4*i+3]: it is read,
 * advanced by exactly the draws the reference's call makes, and written
 * back.  A ray is Ray::new's (src/math.rs:763-773): dir already normalised
 * (inv_dir = 1/dir is formed here), t_min, and t_far = the call's
 * surface_t / max_dist. */
typedef struct {
    double origin[3];
    double dir[3];
    double t_min;
    double t_far;
} aether_medium_ray;
/* MediumInteraction (src/fluid.rs:680-685); hit = 0 for None. */
typedef struct {
    int32_t hit;
    int32_t reserved;
    double t;
    double point[3];
    double albedo[4];   /* scatter_albedo.eval4 at the ray's wavelengths */
    double emission[4]; /* hot_core_emission.eval4 * glow */
} aether_medium_interaction;
/* fluid.rs:793-838 sample_interaction (Woodcock delta tracking).  lambda:
 * 4 wavelengths (nm) per ray, the ray's Wavelengths4. */
int aether_fluid_medium_sample_interaction(aether_fluid_medium_t m, const aether_medium_ray *rays,
                                           size_t nrays, const double *lambda,
                                           uint64_t *rng_state, aether_medium_interaction *out);
/* fluid.rs:841-874 transmittance (ratio tracking, Russian roulette). */
int aether_fluid_medium_transmittance(aether_fluid_medium_t m, const aether_medium_ray *rays,
                                      size_t nrays, uint64_t *rng_state, double *out);
/* fluid.rs:878-888 phase_hg_pdf(wo, wi) for n direction pairs (x,y,z each). */
int aether_fluid_medium_phase_hg_pdf(aether_fluid_medium_t m, const double *wo, const double *wi,
                                     size_t n, double *pdf);
/* fluid.rs:892-910 sample_phase_hg(wo): wi (3 per direction) and its pdf. */
int aether_fluid_medium_sample_phase_hg(aether_fluid_medium_t m, const double *wo, size_t n,
                                        uint64_t *rng_state, double *wi, double *pdf);

/* Seeded inputs: the reference's SplitMix64-seeded Xoshiro256++ stream
 * (src/rng.rs:21-33, 48-64, 74-88).  Fills out[0..n) with range_f64(lo,hi)
 * draws continuing from *state (4 words), which is advanced. */
int aether_rng_seed(uint64_t seed, uint64_t state[4]);
int aether_rng_fill_range(uint64_t state[4], double lo, double hi, double *out, size_t n);

/* Batched instances (no counterpart in the reference, which steps one grid
 * per FluidSolver3D).  Many independent grids of different sizes advance
 * together, one launch per stage over all their tiles.  Shared: dt, solver
 * knobs, dissipation (from `shared`; CG only).  Per instance: grid, h,
 * forcing and a sphere source that sets density/temperature. */
typedef struct aether_fluid_batch_s *aether_fluid_batch_t;
typedef struct {
    int32_t nx, ny, nz;
    int32_t reserved;
    double h;                 /* <= 0: 1/max(nx,ny,nz) as src/fluid.rs:56 */
    double vorticity_eps, buoyancy_beta, gravity_alpha;
    double source_center[3];  /* cell-index coordinates */
    double source_radius;     /* <= 0: no source */
    double source_density, source_temperature;
} aether_fluid_instance_desc;
int aether_fluid_batch_create(const aether_fluid_instance_desc *desc, int count, int precision,
                              const aether_fluid_params *shared, aether_fluid_batch_t *out);
int aether_fluid_batch_destroy(aether_fluid_batch_t b);
int aether_fluid_batch_upload(aether_fluid_batch_t b, int inst, int field, const double *src,
                              size_t n);
int aether_fluid_batch_download(aether_fluid_batch_t b, int inst, int field, double *dst,
                                size_t n);
/* `out` (may be NULL) receives nsteps*count frames, step-major. */
int aether_fluid_batch_run(aether_fluid_batch_t b, double dt, const double *times, int nsteps,
                           aether_fluid_telemetry *out);
/* CG iterations of the last run, nsteps*count values, step-major. */
int aether_fluid_batch_iterations(aether_fluid_batch_t b, int32_t *out, int n, int *count);
/* CG layout and tile autotuner of a batch: times the whole batch's CG in the
 * four-wide layout (FP32, every nx % 4 == 0) with 2 and with 4 z planes per
 * thread, in the one-cell-per-thread layout (8 z cells per thread), and in
 * the one-launch layout solved group by group (FP32, every nx % 4 == 0), and
 * keeps the fastest as the batch's choice for cg_form AETHER_CG_AUTO and
 * cg_z_cells 0 (every instance's pressure and divergence are saved and
 * restored).  iters > 0: `iters` iterations on synthetic right-hand sides,
 * no tolerance stop.  iters == 0: whole solves, from zero pressure, of the
 * last step's right-hand sides at the batch's tolerance and iteration cap,
 * so that instances converging at different iterations count as in the
 * batch's steps (AETHER_EINVAL before the first step).  us_per_iter[4] (may
 * be NULL): time per iteration (iters 0: per iteration of the instance that
 * ran longest) of four-wide z2, four-wide z4, one cell per thread, one launch
 * (-1 when not applicable).  *form / *z_cells: the choice. */
int aether_fluid_batch_autotune_cg(aether_fluid_batch_t b, int iters, double *us_per_iter,
                                   int *form, int *z_cells);

/* Differentiable rollouts (no counterpart in the reference).  Starting from
 * the solver's current state (which is left unchanged), run `nsteps` steps
 * in which the inflow patch [patch_lo, patch_hi) has density and
 * temperature SET to `controls` (patch cells, x fastest) in the source slot
 * of src/fluid.rs:603 and buoyancy_beta = `beta`; return
 * loss = sum over cells (density - target)^2.  If grad_controls / grad_beta
 * are non-NULL, also run the hand-written adjoint (checkpoint every
 * `checkpoint_every` steps, segments recomputed) and return dloss/dcontrols
 * and dloss/dbeta.  The adjoint projection is solved by CG to
 * adjoint_cg_rel_tol (max adjoint_cg_iters). */
typedef struct {
    int32_t nsteps;
    int32_t checkpoint_every;
    double dt;
    int32_t patch_lo[3];
    int32_t patch_hi[3];
    int32_t adjoint_cg_iters;
    int32_t reserved;
    double adjoint_cg_rel_tol;
} aether_rollout_desc;
int aether_fluid_rollout(aether_fluid_t s, const aether_rollout_desc *d, const double *controls,
                         double beta, const double *target, double *loss, double *grad_controls,
                         double *grad_beta);
/* The same rollout as a reverse-mode node (no counterpart in the reference),
 * for a loss the caller defines on the final state.  rollout_forward runs
 * the forward sweep from the current state (left unchanged) and records the
 * tape (checkpoints every checkpoint_every steps, the last segment's
 * intermediates, controls, beta and the knobs); final_state (may be NULL)
 * receives the final u, v, w, density, temperature, each nx*ny*nz f64 in
 * that order.  rollout_backward consumes the tape: lam_final (5*nx*ny*nz,
 * same order; NULL = zero) is the adjoint of the final state, and it
 * returns the vector-Jacobian products for the controls (grad_controls,
 * patch cells), beta (grad_beta) and the initial u, v, w, density,
 * temperature (grad_initial, 5*nx*ny*nz; the patch cells, which the
 * controls overwrite, get 0).  Any argument may be NULL.  The pressure
 * warm start is held fixed (the projection is treated as an exact solve).
 * The solver's state is the rollout's initial state again on return; each
 * forward is consumed by one backward. */
int aether_fluid_rollout_forward(aether_fluid_t s, const aether_rollout_desc *d,
                                 const double *controls, double beta, double *final_state);
int aether_fluid_rollout_backward(aether_fluid_t s, const double *lam_final,
                                  double *grad_controls, double *grad_beta,
                                  double *grad_initial);

/* Measurement hooks used by the benchmark driver. */
/* Average device time (ms) per launch of the CG iteration kernels during the
 * last aether_fluid_run with timing enabled (HIP events on the solver's
 * stream); *launches counts the kernels timed.  which: 0 direction launch
 * (cg_dir), 1 update launch (cg_update), 2 (no longer used: an error),
 * 3 one-launch iteration (cg_fused). */
int aether_fluid_set_kernel_timing(aether_fluid_t s, int enable);
/* Plain device-to-device float4 copy of nbytes (multiple of 16) of
 * full-entropy data, `reps` launches timed with HIP events: the
 * measured-bandwidth reference point. */
int aether_fluid_stream_copy(size_t nbytes, int reps, double *avg_ms);
int aether_fluid_kernel_timing(aether_fluid_t s, int which, double *avg_ms, long *launches);

#ifdef __cplusplus
}
#endif
#endif /* AETHER_FLUID_H */
