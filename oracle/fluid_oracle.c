/*
 * fluid_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, f64 restatement of the reference smoke solver `FluidSolver3D`
 * (src/fluid.rs:23-673) and of its seeded RNG (src/rng.rs:5-88).  It is the
 * parity checker for the HIP library and the CPU baseline timed by bench.py.
 * of_medium_* restate FluidMedium's volumetric sampling (src/fluid.rs:675-910)
 * as the checker of the device sampler.
 * of_soundtrack restates the telemetry-driven audio synthesis and 24-bit
 * dithered encoder (src/audio.rs:1-639) so that telemetry can be checked
 * against the reference's rendered soundtrack; nothing else uses it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * Fidelity rules (so that the f64 results follow the Rust code operation for
 * operation): same loop nests and iteration order, same in-place sequencing,
 * same left-to-right evaluation of every expression, no FMA contraction
 * (build with -ffp-contract=off), sequential reductions in index order.
 * Rust's f64::clamp/min/max are restated as rclamp/fmin/fmax (identical for
 * non-NaN input).
 *
 * Extensions (all default to the reference's behaviour):
 *   h override            (fluid.rs:56 rule when h <= 0)
 *   sor_omega             (fluid.rs:398 hard-codes 1.65)
 *   cg_rel_tol            (stop when ||r||2 <= tol*||b||2; 0 = reference guards only)
 *   dissipation triple    (fluid.rs:599-608 hard-codes 0.996/0.996/0.985)
 *   source_kind           0 = reference jets (fluid.rs:532-595),
 *                         1 = sphere that SETS density/temperature,
 *                         2 = no source
 *   CG statistics         iterations run, ||b||^2, final ||r||^2
 *
 * Parity pins (tests/test_oracle.py): README.md:54 known answer of the
 * verify.rs:229-250 scenario, the fluid.rs:918-968 tests, closed forms
 * (trilinear exactness on affine fields, integer shifts, Laplacian symmetry
 * and null space, brute-force dense Poisson solve, curl of linear fields,
 * telemetry of a single unit mass), and the published Xoshiro256++ /
 * SplitMix64 vectors.
 *
 * Built twice by oracle/Makefile: sequential (the faithful stand-in) and with
 * -fopenmp (the multi-core CPU baseline; its reductions are re-associated by
 * OpenMP, so it is used for timing only, never as the checker).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#define OF_PRAGMA(x) _Pragma(#x)
#define OMP_FOR OF_PRAGMA(omp parallel for schedule(static))
#define OMP_FOR_SUM1(a) OF_PRAGMA(omp parallel for schedule(static) reduction(+ : a))
#else
#define OMP_FOR
#define OMP_FOR_SUM1(a)
#endif

#define TAU_ 6.28318530717958647692528676655900577

typedef struct {
    size_t nx, ny, nz, n;
    double h;
    double *u, *v, *w, *density, *temperature, *pressure, *divergence;
    double *curl_x, *curl_y, *curl_z, *curl_mag;
    double *tmp_fwd, *tmp_bwd, *tmp_out, *snap;
    double *cg_r, *cg_d, *cg_q;
    double vorticity_eps, buoyancy_beta, gravity_alpha;
    int solver_kind; /* 0 = CG, 1 = red-black Gauss-Seidel/SOR */
    long solver_iters;
    double sor_omega, cg_rel_tol;
    double diss_vel, diss_rho, diss_temp;
    int source_kind;
    double src_cx, src_cy, src_cz, src_r, src_rho, src_temp;
    /* statistics of the last pressure solve */
    long stat_iters;
    double stat_b2, stat_r2_init, stat_r2;
    int stat_converged;
} OFluid;

/* Field ids shared with the product ABI. */
enum { F_U = 0, F_V, F_W, F_DENSITY, F_TEMPERATURE, F_PRESSURE, F_DIVERGENCE,
       F_CURL_MAG, F_CG_R };

static inline size_t IX(const OFluid *s, size_t i, size_t j, size_t k) {
    return (k * s->ny + j) * s->nx + i; /* fluid.rs:88-90 */
}

static inline double rclamp(double x, double lo, double hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */

/* fluid.rs:53-85 (h_override > 0 replaces the 1/max(n) rule of fluid.rs:56) */
OFluid *of_create(long nx, long ny, long nz, double h_override) {
    if (nx < 6 || ny < 6 || nz < 6) return NULL; /* fluid.rs:54 */
    OFluid *s = (OFluid *)calloc(1, sizeof(OFluid));
    s->nx = (size_t)nx; s->ny = (size_t)ny; s->nz = (size_t)nz;
    s->n = s->nx * s->ny * s->nz;
    size_t m = s->nx;
    if (s->ny > m) m = s->ny;
    if (s->nz > m) m = s->nz;
    s->h = h_override > 0.0 ? h_override : 1.0 / (double)m;
    double **arrs[] = {&s->u, &s->v, &s->w, &s->density, &s->temperature, &s->pressure,
                       &s->divergence, &s->curl_x, &s->curl_y, &s->curl_z, &s->curl_mag,
                       &s->tmp_fwd, &s->tmp_bwd, &s->tmp_out, &s->snap, &s->cg_r, &s->cg_d,
                       &s->cg_q};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a)
        *arrs[a] = (double *)calloc(s->n, sizeof(double));
    s->vorticity_eps = 5.5; /* fluid.rs:79-83 */
    s->buoyancy_beta = 6.8;
    s->gravity_alpha = 0.45;
    s->solver_kind = 0;
    s->solver_iters = 32;
    s->sor_omega = 1.65;  /* fluid.rs:398 */
    s->cg_rel_tol = 0.0;
    s->diss_vel = 0.996;  /* fluid.rs:599-607 */
    s->diss_rho = 0.996;
    s->diss_temp = 0.985; /* fluid.rs:608 */
    s->source_kind = 0;
    return s;
}

void of_destroy(OFluid *s) {
    if (!s) return;
    double *arrs[] = {s->u, s->v, s->w, s->density, s->temperature, s->pressure,
                      s->divergence, s->curl_x, s->curl_y, s->curl_z, s->curl_mag,
                      s->tmp_fwd, s->tmp_bwd, s->tmp_out, s->snap, s->cg_r, s->cg_d, s->cg_q};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) free(arrs[a]);
    free(s);
}

void of_set_params(OFluid *s, double eps, double beta, double alpha, int solver_kind,
                   long solver_iters, double omega, double rel_tol, double diss_vel,
                   double diss_rho, double diss_temp) {
    s->vorticity_eps = eps;
    s->buoyancy_beta = beta;
    s->gravity_alpha = alpha;
    s->solver_kind = solver_kind;
    s->solver_iters = solver_iters;
    s->sor_omega = omega;
    s->cg_rel_tol = rel_tol;
    s->diss_vel = diss_vel;
    s->diss_rho = diss_rho;
    s->diss_temp = diss_temp;
}

void of_set_source(OFluid *s, int kind, double cx, double cy, double cz, double r,
                   double rho, double temp) {
    s->source_kind = kind;
    s->src_cx = cx; s->src_cy = cy; s->src_cz = cz;
    s->src_r = r; s->src_rho = rho; s->src_temp = temp;
}

double *of_field(OFluid *s, int id) {
    switch (id) {
    case F_U: return s->u;
    case F_V: return s->v;
    case F_W: return s->w;
    case F_DENSITY: return s->density;
    case F_TEMPERATURE: return s->temperature;
    case F_PRESSURE: return s->pressure;
    case F_DIVERGENCE: return s->divergence;
    case F_CURL_MAG: return s->curl_mag;
    case F_CG_R: return s->cg_r;
    default: return NULL;
    }
}

double of_h(const OFluid *s) { return s->h; }

void of_stats(const OFluid *s, long *iters, double *b2, double *r2_init, double *r2,
              int *converged) {
    *iters = s->stat_iters;
    *b2 = s->stat_b2;
    *r2_init = s->stat_r2_init;
    *r2 = s->stat_r2;
    *converged = s->stat_converged;
}

/* ------------------------------------------------------------------ */
/* sampling and advection                                              */
/* ------------------------------------------------------------------ */

/* fluid.rs:94-128 */
static double sample_trilinear(const OFluid *s, const double *f, double x, double y, double z) {
    double xc = rclamp(x, 0.5, (double)s->nx - 1.500001);
    double yc = rclamp(y, 0.5, (double)s->ny - 1.500001);
    double zc = rclamp(z, 0.5, (double)s->nz - 1.500001);
    size_t i0 = (size_t)floor(xc), j0 = (size_t)floor(yc), k0 = (size_t)floor(zc);
    size_t i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
    double tx = xc - (double)i0, ty = yc - (double)j0, tz = zc - (double)k0;
    double c000 = f[IX(s, i0, j0, k0)], c100 = f[IX(s, i1, j0, k0)];
    double c010 = f[IX(s, i0, j1, k0)], c110 = f[IX(s, i1, j1, k0)];
    double c001 = f[IX(s, i0, j0, k1)], c101 = f[IX(s, i1, j0, k1)];
    double c011 = f[IX(s, i0, j1, k1)], c111 = f[IX(s, i1, j1, k1)];
    double c00 = c000 * (1.0 - tx) + c100 * tx;
    double c10 = c010 * (1.0 - tx) + c110 * tx;
    double c01 = c001 * (1.0 - tx) + c101 * tx;
    double c11 = c011 * (1.0 - tx) + c111 * tx;
    double c0 = c00 * (1.0 - ty) + c10 * ty;
    double c1 = c01 * (1.0 - ty) + c11 * ty;
    return c0 * (1.0 - tz) + c1 * tz;
}

double of_sample_trilinear(const OFluid *s, int field, double x, double y, double z) {
    return sample_trilinear(s, of_field((OFluid *)s, field), x, y, z);
}

/* FluidMedium (fluid.rs:687-788), the renderer-side consumer of the
 * solver's density and temperature. */

/* fluid.rs:704-708: max over the density array, starting from 0 */
double of_max_density(const OFluid *s) {
    double m = 0.0;
    for (size_t a = 0; a < s->n; ++a) m = fmax(m, s->density[a]);
    return m;
}

/* fluid.rs:739-788 sample_density_and_temp, bounds = [bmin, bmax]
 * (Aabb::extent = max(max - min, 0) per axis, math.rs:832-834) */
void of_medium_sample(const OFluid *s, const double *bmin, const double *bmax, double px,
                      double py, double pz, double *d_out, double *t_out) {
    if (px < bmin[0] || px > bmax[0] || py < bmin[1] || py > bmax[1] || pz < bmin[2] ||
        pz > bmax[2]) {
        *d_out = 0.0;
        *t_out = 0.0;
        return;
    }
    double ex = fmax(bmax[0] - bmin[0], 0.0), ey = fmax(bmax[1] - bmin[1], 0.0),
           ez = fmax(bmax[2] - bmin[2], 0.0);
    double ux = (px - bmin[0]) / fmax(ex, 1e-9);
    double uy = (py - bmin[1]) / fmax(ey, 1e-9);
    double uz = (pz - bmin[2]) / fmax(ez, 1e-9);
    double fx = rclamp(4.0 * ux * (1.0 - ux), 0.0, 1.0);
    double fz = rclamp(4.0 * uz * (1.0 - uz), 0.0, 1.0);
    double fy = rclamp(1.0 - fmax(uy - 0.95, 0.0) / 0.05, 0.0, 1.0);
    double feather = sqrt(fx * fz * fy);
    double gx = rclamp(ux * (double)(s->nx - 1), 0.5, (double)s->nx - 1.500001);
    double gy = rclamp(uy * (double)(s->ny - 1), 0.5, (double)s->ny - 1.500001);
    double gz = rclamp(uz * (double)(s->nz - 1), 0.5, (double)s->nz - 1.500001);
    /* the clamped position lies in the clamp box, where sample_trilinear's
     * own clamp is the identity: the same interpolation, same association */
    *d_out = sample_trilinear(s, s->density, gx, gy, gz) * feather;
    *t_out = sample_trilinear(s, s->temperature, gx, gy, gz) * feather;
}

/* fluid.rs:131-150 */
static void stencil_minmax(const OFluid *s, const double *f, double x, double y, double z,
                           double *lo_out, double *hi_out) {
    double xc = rclamp(x, 0.5, (double)s->nx - 1.500001);
    double yc = rclamp(y, 0.5, (double)s->ny - 1.500001);
    double zc = rclamp(z, 0.5, (double)s->nz - 1.500001);
    size_t i0 = (size_t)floor(xc), j0 = (size_t)floor(yc), k0 = (size_t)floor(zc);
    double lo = INFINITY, hi = -INFINITY;
    for (size_t dk = 0; dk <= 1; ++dk)
        for (size_t dj = 0; dj <= 1; ++dj)
            for (size_t di = 0; di <= 1; ++di) {
                double v = f[IX(s, i0 + di, j0 + dj, k0 + dk)];
                lo = fmin(lo, v);
                hi = fmax(hi, v);
            }
    *lo_out = lo;
    *hi_out = hi;
}

/* fluid.rs:153-167 */
static void advect_semi_lagrangian(const OFluid *s, const double *src, double *dst, double dt) {
    double scale = dt / s->h;
    memset(dst, 0, s->n * sizeof(double));
    OMP_FOR
    for (size_t k = 1; k < s->nz - 1; ++k)
        for (size_t j = 1; j < s->ny - 1; ++j)
            for (size_t i = 1; i < s->nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double bx = (double)i - scale * s->u[id];
                double by = (double)j - scale * s->v[id];
                double bz = (double)k - scale * s->w[id];
                dst[id] = sample_trilinear(s, src, bx, by, bz);
            }
}

/* fluid.rs:170-234 */
void of_advect(OFluid *s, int kind, double dt, double dissipation) {
    double *field = kind == 0 ? s->u : kind == 1 ? s->v : kind == 2 ? s->w
                  : kind == 3 ? s->density : s->temperature;
    const double *src = s->snap;
    memcpy(s->snap, field, s->n * sizeof(double));
    advect_semi_lagrangian(s, src, s->tmp_fwd, dt);
    advect_semi_lagrangian(s, s->tmp_fwd, s->tmp_bwd, -dt);
    double *out = s->tmp_out;
    memset(out, 0, s->n * sizeof(double));
    double scale = dt / s->h;
    OMP_FOR
    for (size_t k = 1; k < s->nz - 1; ++k)
        for (size_t j = 1; j < s->ny - 1; ++j)
            for (size_t i = 1; i < s->nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double bx = (double)i - scale * s->u[id];
                double by = (double)j - scale * s->v[id];
                double bz = (double)k - scale * s->w[id];
                int near_wall = bx < 1.5 || bx > (double)s->nx - 2.5 || by < 1.5 ||
                                by > (double)s->ny - 2.5 || bz < 1.5 || bz > (double)s->nz - 2.5;
                double val;
                if (near_wall) {
                    val = s->tmp_fwd[id];
                } else {
                    double mc = s->tmp_fwd[id] + 0.5 * (src[id] - s->tmp_bwd[id]);
                    double lo, hi;
                    stencil_minmax(s, src, bx, by, bz, &lo, &hi);
                    val = rclamp(mc, lo, hi);
                }
                out[id] = val * dissipation;
            }
    if (kind <= 2) {
        memcpy(field, out, s->n * sizeof(double));
    } else {
        OMP_FOR
        for (size_t a = 0; a < s->n; ++a) field[a] = fmax(out[a], 0.0);
    }
}

/* fluid.rs:236-253 */
void of_enforce_bc(OFluid *s) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    OMP_FOR
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                size_t id = IX(s, i, j, k);
                if (i <= 1 || i + 1 >= nx || j == 0 || j + 1 >= ny || k == 0 || k + 1 >= nz)
                    s->u[id] = 0.0;
                if (j <= 1 || j + 1 >= ny || i == 0 || i + 1 >= nx || k == 0 || k + 1 >= nz)
                    s->v[id] = 0.0;
                if (k <= 1 || k + 1 >= nz || i == 0 || i + 1 >= nx || j == 0 || j + 1 >= ny)
                    s->w[id] = 0.0;
            }
}

/* fluid.rs:256-330 */
void of_buoyancy_confinement(OFluid *s, double dt) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    double inv_2h = 0.5 / s->h;
    memset(s->curl_mag, 0, s->n * sizeof(double));
    OMP_FOR
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double dw_dy = (s->w[IX(s, i, j + 1, k)] - s->w[IX(s, i, j - 1, k)]) * inv_2h;
                double dv_dz = (s->v[IX(s, i, j, k + 1)] - s->v[IX(s, i, j, k - 1)]) * inv_2h;
                double du_dz = (s->u[IX(s, i, j, k + 1)] - s->u[IX(s, i, j, k - 1)]) * inv_2h;
                double dw_dx = (s->w[IX(s, i + 1, j, k)] - s->w[IX(s, i - 1, j, k)]) * inv_2h;
                double dv_dx = (s->v[IX(s, i + 1, j, k)] - s->v[IX(s, i - 1, j, k)]) * inv_2h;
                double du_dy = (s->u[IX(s, i, j + 1, k)] - s->u[IX(s, i, j - 1, k)]) * inv_2h;
                double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
                s->curl_x[id] = ox;
                s->curl_y[id] = oy;
                s->curl_z[id] = oz;
                s->curl_mag[id] = sqrt(ox * ox + oy * oy + oz * oz);
            }
    double eps_h = s->vorticity_eps * s->h;
    OMP_FOR
    for (size_t k = 2; k < nz - 2; ++k)
        for (size_t j = 2; j < ny - 2; ++j)
            for (size_t i = 2; i < nx - 2; ++i) {
                size_t id = IX(s, i, j, k);
                const double *cm = s->curl_mag;
                double gx = (cm[IX(s, i + 1, j, k)] - cm[IX(s, i - 1, j, k)]) * inv_2h;
                double gy = (cm[IX(s, i, j + 1, k)] - cm[IX(s, i, j - 1, k)]) * inv_2h;
                double gz = (cm[IX(s, i, j, k + 1)] - cm[IX(s, i, j, k - 1)]) * inv_2h;
                double glen = sqrt(gx * gx + gy * gy + gz * gz) + 1e-10;
                double nxv = gx / glen, nyv = gy / glen, nzv = gz / glen;
                double ox = s->curl_x[id], oy = s->curl_y[id], oz = s->curl_z[id];
                double fx = eps_h * (nyv * oz - nzv * oy);
                double fy = eps_h * (nzv * ox - nxv * oz);
                double fz = eps_h * (nxv * oy - nyv * ox);
                s->u[id] += dt * fx;
                s->v[id] += dt * fy;
                s->w[id] += dt * fz;
            }
    OMP_FOR
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double lift = s->buoyancy_beta * s->temperature[id] -
                              s->gravity_alpha * s->density[id];
                s->v[id] += dt * lift;
            }
    of_enforce_bc(s);
}

/* fluid.rs:333-353 */
double of_divergence_norm(OFluid *s) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    double inv_h = 1.0 / s->h;
    double sum_sq = 0.0;
    memset(s->divergence, 0, s->n * sizeof(double));
    OMP_FOR_SUM1(sum_sq)
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double du = s->u[IX(s, i + 1, j, k)] - s->u[id];
                double dv = s->v[IX(s, i, j + 1, k)] - s->v[id];
                double dw = s->w[IX(s, i, j, k + 1)] - s->w[id];
                double div_val = (du + dv + dw) * inv_h;
                s->divergence[id] = -s->h * (du + dv + dw);
                sum_sq += div_val * div_val;
            }
    double interior_n = (double)((nx - 2) * (ny - 2) * (nz - 2));
    return sqrt(sum_sq / interior_n);
}

/* Extension: max |div u| over the interior (same stencil as fluid.rs:341-344). */
double of_divergence_max(const OFluid *s) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    double inv_h = 1.0 / s->h, m = 0.0;
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double du = s->u[IX(s, i + 1, j, k)] - s->u[id];
                double dv = s->v[IX(s, i, j + 1, k)] - s->v[id];
                double dw = s->w[IX(s, i, j, k + 1)] - s->w[id];
                m = fmax(m, fabs((du + dv + dw) * inv_h));
            }
    return m;
}

/* fluid.rs:356-394 */
static void apply_spd_laplacian(const OFluid *s, const double *x, double *out) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    memset(out, 0, s->n * sizeof(double));
    OMP_FOR
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double diag = 0.0, nb_sum = 0.0;
                if (i > 1) { diag += 1.0; nb_sum += x[IX(s, i - 1, j, k)]; }
                if (i + 2 < nx) { diag += 1.0; nb_sum += x[IX(s, i + 1, j, k)]; }
                if (j > 1) { diag += 1.0; nb_sum += x[IX(s, i, j - 1, k)]; }
                if (j + 2 < ny) { diag += 1.0; nb_sum += x[IX(s, i, j + 1, k)]; }
                if (k > 1) { diag += 1.0; nb_sum += x[IX(s, i, j, k - 1)]; }
                if (k + 2 < nz) { diag += 1.0; nb_sum += x[IX(s, i, j, k + 1)]; }
                out[id] = diag * x[id] - nb_sum;
            }
}

/* Exposed for the operator tests: out = A x with A the fluid.rs:356-394 matrix. */
void of_apply_laplacian(const OFluid *s, const double *x, double *out) {
    apply_spd_laplacian(s, x, out);
}

/* fluid.rs:397-442 (omega parameterised; the reference constant is 1.65) */
void of_solve_rbgs(OFluid *s, long iterations) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    double omega = s->sor_omega;
    double *p = s->pressure;
    const double *b = s->divergence;
    for (long it = 0; it < iterations; ++it)
        for (size_t color = 0; color < 2; ++color) {
            OMP_FOR
            for (size_t k = 1; k < nz - 1; ++k)
                for (size_t j = 1; j < ny - 1; ++j)
                    for (size_t i = 1; i < nx - 1; ++i) {
                        if ((i + j + k) % 2 != color) continue;
                        size_t id = IX(s, i, j, k);
                        double diag = 0.0, nb_sum = 0.0;
                        if (i > 1) { diag += 1.0; nb_sum += p[IX(s, i - 1, j, k)]; }
                        if (i + 2 < nx) { diag += 1.0; nb_sum += p[IX(s, i + 1, j, k)]; }
                        if (j > 1) { diag += 1.0; nb_sum += p[IX(s, i, j - 1, k)]; }
                        if (j + 2 < ny) { diag += 1.0; nb_sum += p[IX(s, i, j + 1, k)]; }
                        if (k > 1) { diag += 1.0; nb_sum += p[IX(s, i, j, k - 1)]; }
                        if (k + 2 < nz) { diag += 1.0; nb_sum += p[IX(s, i, j, k + 1)]; }
                        double gs_target = (b[id] + nb_sum) / diag;
                        p[id] = (1.0 - omega) * p[id] + omega * gs_target;
                    }
        }
    s->stat_iters = iterations;
    s->stat_converged = 1;
}


/* fluid.rs:445-493, plus an iteration counter and the optional relative
 * tolerance ||r||2 <= tol*||b||2, tested at the same points as the
 * reference's absolute guards (fluid.rs:462, 479). */
void of_solve_cg(OFluid *s, long max_iters) {
    size_t n = s->n;
    double *r = s->cg_r, *d = s->cg_d, *q = s->cg_q, *p = s->pressure;
    const double *b = s->divergence;
    apply_spd_laplacian(s, p, q); /* fluid.rs:450-453 warm start */
    double delta_new = 0.0;
    for (size_t a = 0; a < n; ++a) {
        double ri = b[a] - q[a];
        r[a] = ri;
        d[a] = ri;
        delta_new += ri * ri;
    }
    double b2 = 0.0;
    for (size_t a = 0; a < n; ++a) b2 += b[a] * b[a];
    double tol2 = s->cg_rel_tol * s->cg_rel_tol * b2;
    int use_rel = s->cg_rel_tol > 0.0;
    long iters = 0;
    int converged = 0;
    s->stat_b2 = b2;
    s->stat_r2_init = delta_new;
    if (delta_new > 1e-24 && !(use_rel && delta_new <= tol2)) {
        for (long it = 0; it < max_iters; ++it) {
            apply_spd_laplacian(s, d, q);
            double dq = 0.0;
            OMP_FOR_SUM1(dq)
            for (size_t a = 0; a < n; ++a) dq += d[a] * q[a];
            if (fabs(dq) < 1e-28) { converged = 1; break; }
            double alpha = delta_new / dq;
            double delta_next = 0.0;
            OMP_FOR_SUM1(delta_next)
            for (size_t a = 0; a < n; ++a) {
                p[a] += alpha * d[a];
                r[a] -= alpha * q[a];
                delta_next += r[a] * r[a];
            }
            ++iters;
            if (delta_next < 1e-20 || (use_rel && delta_next <= tol2)) {
                delta_new = delta_next;
                converged = 1;
                break;
            }
            double beta = delta_next / delta_new;
            OMP_FOR
            for (size_t a = 0; a < n; ++a) d[a] = r[a] + beta * d[a];
            delta_new = delta_next;
        }
    } else {
        converged = 1;
    }
    s->stat_iters = iters;
    s->stat_r2 = delta_new;
    s->stat_converged = converged;
}

/* fluid.rs:496-529 */
void of_project(OFluid *s) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    of_enforce_bc(s);
    of_divergence_norm(s);
    if (s->solver_kind == 0) of_solve_cg(s, s->solver_iters);
    else of_solve_rbgs(s, s->solver_iters);
    double inv_h = 1.0 / s->h;
    const double *p = s->pressure;
    OMP_FOR
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double pc = p[id];
                if (i > 1) s->u[id] -= (pc - p[IX(s, i - 1, j, k)]) * inv_h;
                if (j > 1) s->v[id] -= (pc - p[IX(s, i, j - 1, k)]) * inv_h;
                if (k > 1) s->w[id] -= (pc - p[IX(s, i, j, k - 1)]) * inv_h;
            }
    of_enforce_bc(s);
}

/* fluid.rs:532-595 */
static void inject_jets(OFluid *s, double time_sec, double dt) {
    double cx = 0.5 * (double)s->nx;
    double cz = 0.5 * (double)s->nz;
    size_t base_j = 3 < s->ny - 3 ? 3 : s->ny - 3;
    double pulse = 0.65 + 0.35 * sin(time_sec * TAU_ * 0.55) +
                   0.25 * fmax(cos(time_sec * TAU_ * 1.35), 0.0);
    double angle = time_sec * 1.85;
    double orbit_r = (double)s->nx * 0.14;
    double jets[2][6] = {
        {cx + orbit_r * cos(angle), (double)base_j + 1.5, cz + orbit_r * sin(angle),
         -2.6 * sin(angle), 7.2 * pulse, 2.6 * cos(angle)},
        {cx - orbit_r * cos(angle), (double)base_j + 2.0, cz - orbit_r * sin(angle),
         2.6 * sin(angle), 6.4 * pulse, -2.6 * cos(angle)},
    };
    double r_cells = fmax((double)s->nx * 0.105, 2.2);
    double r2 = r_cells * r_cells;
    for (int q = 0; q < 2; ++q) {
        double jx = jets[q][0], jy = jets[q][1], jz = jets[q][2];
        double vx = jets[q][3], vy = jets[q][4], vz = jets[q][5];
        size_t k_min = (size_t)fmax(floor(jz - r_cells - 1.0), 1.0);
        size_t k_max = (size_t)fmin(ceil(jz + r_cells + 1.0), (double)(s->nz - 2));
        size_t j_min = (size_t)fmax(floor(jy - r_cells - 1.0), 1.0);
        size_t j_max = (size_t)fmin(ceil(jy + r_cells + 1.0), (double)(s->ny - 2));
        size_t i_min = (size_t)fmax(floor(jx - r_cells - 1.0), 1.0);
        size_t i_max = (size_t)fmin(ceil(jx + r_cells + 1.0), (double)(s->nx - 2));
        for (size_t k = k_min; k <= k_max; ++k)
            for (size_t j = j_min; j <= j_max; ++j)
                for (size_t i = i_min; i <= i_max; ++i) {
                    double dx = (double)i - jx, dy = (double)j - jy, dz = (double)k - jz;
                    double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < r2) {
                        double w_falloff = exp(-2.2 * d2 / r2);
                        size_t id = IX(s, i, j, k);
                        s->density[id] = fmin(s->density[id] + dt * 9.5 * pulse * w_falloff, 3.8);
                        s->temperature[id] =
                            fmin(s->temperature[id] + dt * 8.2 * pulse * w_falloff, 4.5);
                        s->u[id] += dt * vx * w_falloff * 4.2;
                        s->v[id] += dt * vy * w_falloff * 4.2;
                        s->w[id] += dt * vz * w_falloff * 4.2;
                    }
                }
    }
}

/* Extension: a sphere (cell-index coordinates, d^2 < r^2) whose interior
 * cells [1, n-2] get density and temperature SET to fixed values. */
static void inject_sphere(OFluid *s) {
    double r2 = s->src_r * s->src_r;
    for (size_t k = 1; k < s->nz - 1; ++k)
        for (size_t j = 1; j < s->ny - 1; ++j)
            for (size_t i = 1; i < s->nx - 1; ++i) {
                double dx = (double)i - s->src_cx, dy = (double)j - s->src_cy,
                       dz = (double)k - s->src_cz;
                if (dx * dx + dy * dy + dz * dz < r2) {
                    size_t id = IX(s, i, j, k);
                    s->density[id] = s->src_rho;
                    s->temperature[id] = s->src_temp;
                }
            }
}

void of_inject_sources(OFluid *s, double time_sec, double dt) {
    if (s->source_kind == 0) inject_jets(s, time_sec, dt);
    else if (s->source_kind == 1) inject_sphere(s);
}

/* fluid.rs:613-672.  out = {time, KE, enstrophy, peak |w|, mass, com x, y, z} */
void of_telemetry(const OFluid *s, double time_sec, double *out) {
    size_t nx = s->nx, ny = s->ny, nz = s->nz;
    double dv = s->h * s->h * s->h;
    double inv_2h = 0.5 / s->h;
    double ke = 0.0, enstrophy = 0.0, peak = 0.0, mass = 0.0;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    const double *u = s->u, *v = s->v, *w = s->w;
    for (size_t k = 1; k < nz - 1; ++k)
        for (size_t j = 1; j < ny - 1; ++j)
            for (size_t i = 1; i < nx - 1; ++i) {
                size_t id = IX(s, i, j, k);
                double u2 = u[id] * u[id] + v[id] * v[id] + w[id] * w[id];
                ke += 0.5 * u2 * dv;
                double dw_dy = (w[IX(s, i, j + 1, k)] - w[IX(s, i, j - 1, k)]) * inv_2h;
                double dv_dz = (v[IX(s, i, j, k + 1)] - v[IX(s, i, j, k - 1)]) * inv_2h;
                double du_dz = (u[IX(s, i, j, k + 1)] - u[IX(s, i, j, k - 1)]) * inv_2h;
                double dw_dx = (w[IX(s, i + 1, j, k)] - w[IX(s, i - 1, j, k)]) * inv_2h;
                double dv_dx = (v[IX(s, i + 1, j, k)] - v[IX(s, i - 1, j, k)]) * inv_2h;
                double du_dy = (u[IX(s, i, j + 1, k)] - u[IX(s, i, j - 1, k)]) * inv_2h;
                double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
                double w2 = ox * ox + oy * oy + oz * oz;
                enstrophy += 0.5 * w2 * dv;
                peak = fmax(peak, sqrt(w2));
                double rho = s->density[id];
                if (rho > 1e-8) {
                    mass += rho * dv;
                    double nxn = 2.0 * ((double)i / (double)(nx - 1)) - 1.0;
                    double nyn = 2.0 * ((double)j / (double)(ny - 1)) - 1.0;
                    double nzn = 2.0 * ((double)k / (double)(nz - 1)) - 1.0;
                    double m = rho * dv;
                    cx += nxn * m; /* Vec3 * f64 then += (math.rs:167-210) */
                    cy += nyn * m;
                    cz += nzn * m;
                }
            }
    if (mass > 1e-12) { /* Vec3 /= f64 multiplies by 1/rhs (math.rs:226-239) */
        double inv = 1.0 / mass;
        cx = cx * inv; cy = cy * inv; cz = cz * inv;
    }
    out[0] = time_sec;
    out[1] = ke;
    out[2] = enstrophy;
    out[3] = peak;
    out[4] = mass;
    out[5] = rclamp(cx, -1.0, 1.0);
    out[6] = rclamp(cy, -1.0, 1.0);
    out[7] = rclamp(cz, -1.0, 1.0);
}

/* fluid.rs:598-611 (dissipation factors parameterised) */
void of_step(OFluid *s, double dt, double time_sec, double *telemetry_out) {
    of_advect(s, 0, dt, s->diss_vel);
    of_advect(s, 1, dt, s->diss_vel);
    of_advect(s, 2, dt, s->diss_vel);
    of_inject_sources(s, time_sec, dt);
    of_buoyancy_confinement(s, dt);
    of_project(s);
    of_advect(s, 3, dt, s->diss_rho);
    of_advect(s, 4, dt, s->diss_temp);
    if (telemetry_out) of_telemetry(s, time_sec, telemetry_out);
}

int of_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* seeded RNG: SplitMix64 -> Xoshiro256++ (rng.rs:5-88)                */
/* ------------------------------------------------------------------ */

typedef struct { uint64_t s[4]; } ORng;

static uint64_t splitmix64(uint64_t *state) { /* rng.rs:5-11 */
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void orng_seed(ORng *r, uint64_t seed) { /* rng.rs:21-33 */
    uint64_t sm = seed ^ 0xA5A55A5AF0F00F0Full;
    for (int a = 0; a < 4; ++a) r->s[a] = splitmix64(&sm);
    if (!r->s[0] && !r->s[1] && !r->s[2] && !r->s[3]) r->s[0] = 1;
}

void orng_set_state(ORng *r, const uint64_t *st) { memcpy(r->s, st, sizeof(r->s)); }

uint64_t orng_next_u64(ORng *r) { /* rng.rs:48-64 */
    uint64_t result = rotl64(r->s[0] + r->s[3], 23) + r->s[0];
    uint64_t t = r->s[1] << 17;
    r->s[2] ^= r->s[0];
    r->s[3] ^= r->s[1];
    r->s[1] ^= r->s[2];
    r->s[0] ^= r->s[3];
    r->s[2] ^= t;
    r->s[3] = rotl64(r->s[3], 45);
    return result;
}

double orng_next_f64(ORng *r) { /* rng.rs:74-77 */
    return (double)(orng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}

double orng_range_f64(ORng *r, double lo, double hi) { /* rng.rs:86-88 */
    return lo + (hi - lo) * orng_next_f64(r);
}

uint64_t of_splitmix64_next(uint64_t *state) { return splitmix64(state); }

/* ------------------------------------------------------------------ */
/* FluidMedium volumetric sampling (fluid.rs:793-910), with the math.rs  */
/* pieces it calls: Vec3::dot/normalize (math.rs:34-65, dot is fused     */
/* multiply-adds), Ray::at (math.rs:776-778), Aabb::intersect            */
/* (math.rs:859-893), Onb::from_normal/to_world (math.rs:727-740),       */
/* SpectralCurve::eval_at (math.rs:1148-1170), Rng::uniform_sphere       */
/* (rng.rs:133-138).  The medium is the oracle's current density and     */
/* temperature (FluidMedium::from_solver clones them, fluid.rs:703-735). */
/* ------------------------------------------------------------------ */

typedef struct {
    double lo[3], hi[3];   /* bounds */
    double sigma_t_scale;
    double max_density;
    double majorant;       /* max(max_density * sigma_t_scale, 1e-4) */
    double phase_g;        /* clamped to [-0.9, 0.9] */
} OMedium;

/* one ray of Ray::new: dir normalised, inv_dir = 1 / dir per axis */
typedef struct {
    double o[3], d[3], inv[3], t_min;
} ORay;

static const double O_PI = 3.14159265358979323846264338327950288;
static const double O_TAU = 6.28318530717958647692528676655900577;

static double vdot(const double *a, const double *b) { /* math.rs:34-36 */
    return fma(a[0], b[0], fma(a[1], b[1], a[2] * b[2]));
}

static void vnormalize(const double *v, double *out) { /* math.rs:58-65 */
    const double len = sqrt(vdot(v, v));
    if (len > 1e-18) {
        const double inv = 1.0 / len;
        out[0] = v[0] * inv;
        out[1] = v[1] * inv;
        out[2] = v[2] * inv;
    } else {
        out[0] = 0.0;
        out[1] = 0.0;
        out[2] = 1.0;
    }
}

void of_ray_make(const double *o, const double *d, double t_min, ORay *r) { /* math.rs:763-773 */
    double n[3];
    vnormalize(d, n);
    for (int a = 0; a < 3; ++a) {
        r->o[a] = o[a];
        r->d[a] = n[a];
        r->inv[a] = 1.0 / n[a];
    }
    r->t_min = t_min;
}

/* math.rs:859-893; f64::min/max ignore a NaN operand like fmin/fmax */
static int aabb_intersect(const OMedium *m, const ORay *r, double t_min, double t_max,
                          double *t0, double *t1) {
    for (int a = 0; a < 3; ++a) {
        const double ta = (m->lo[a] - r->o[a]) * r->inv[a];
        const double tb = (m->hi[a] - r->o[a]) * r->inv[a];
        const double lo = fmin(ta, tb), hi = fmax(ta, tb);
        t_min = fmax(t_min, lo);
        t_max = fmin(t_max, hi);
        if (a < 2 && t_max < t_min) return 0;
    }
    if (t_max >= t_min) {
        *t0 = t_min;
        *t1 = t_max;
        return 1;
    }
    return 0;
}

/* SpectralCurve::eval_at (math.rs:1148-1170): kind 0 = SmoothRgb(c),
 * 1 = NeonDischarge with FluidMedium's constants (fluid.rs:723-731) */
static double spectral_eval(int kind, double lambda) {
    const double w_b = 1.0 / (1.0 + exp((lambda - 492.0) / 14.0));
    const double w_r = 1.0 / (1.0 + exp(-(lambda - 586.0) / 14.0));
    const double w_g = fmax(1.0 - w_b - w_r, 0.0);
    if (kind == 0) return 0.84 * w_r + 0.90 * w_g + 0.96 * w_b; /* fluid.rs:721 */
    const double base = 0.18 * w_r + 0.72 * w_g + 1.15 * w_b;
    const double e1 = (lambda - 486.0) / 14.0, e2 = (lambda - 618.0) / 16.0;
    const double g1 = 1.6 * exp(-0.5 * (e1 * e1));
    const double g2 = 1.1 * exp(-0.5 * (e2 * e2));
    return base + g1 + g2;
}

/* fluid.rs:739-788 at p, for the medium's own bounds */
static void medium_at(const OFluid *s, const OMedium *m, const double *p, double *rho,
                      double *temp) {
    of_medium_sample(s, m->lo, m->hi, p[0], p[1], p[2], rho, temp);
}

/* fluid.rs:793-838 sample_interaction.  out: [hit, t, px, py, pz, albedo x4,
 * emission x4] (14 doubles); the RNG is advanced as the reference's. */
void of_medium_sample_interaction(const OFluid *s, const OMedium *m, const ORay *r,
                                  double surface_t, const double *lambda, ORng *rng,
                                  double *out) {
    for (int a = 0; a < 14; ++a) out[a] = 0.0;
    if (m->max_density <= 1e-6) return;
    double t_enter, t_exit;
    if (!aabb_intersect(m, r, r->t_min, surface_t, &t_enter, &t_exit)) return;
    double t = fmax(t_enter, r->t_min);
    const double limit = fmin(t_exit, surface_t);
    if (t >= limit) return;
    const double inv_maj = 1.0 / m->majorant;
    for (int it = 0; it < 256; ++it) {
        const double u = fmax(1.0 - orng_next_f64(rng), 1e-15);
        t -= log(u) * inv_maj;
        if (t >= limit) return;
        const double p[3] = {r->o[0] + r->d[0] * t, r->o[1] + r->d[1] * t,
                             r->o[2] + r->d[2] * t};
        double rho, temp;
        medium_at(s, m, p, &rho, &temp);
        const double sigma_t = rho * m->sigma_t_scale;
        if (orng_next_f64(rng) * m->majorant < sigma_t) {
            out[0] = 1.0;
            out[1] = t;
            out[2] = p[0];
            out[3] = p[1];
            out[4] = p[2];
            const double gb = fmax(temp - 0.35, 0.0);
            const double glow = gb * gb * 0.95;
            for (int w = 0; w < 4; ++w) {
                out[5 + w] = spectral_eval(0, lambda[w]);
                out[9 + w] = glow > 1e-5 ? spectral_eval(1, lambda[w]) * glow : 0.0;
            }
            return;
        }
    }
}

/* fluid.rs:841-874 transmittance (ratio tracking with Russian roulette) */
double of_medium_transmittance(const OFluid *s, const OMedium *m, const ORay *r, double max_dist,
                               ORng *rng) {
    if (m->max_density <= 1e-6) return 1.0;
    double t_enter, t_exit;
    if (!aabb_intersect(m, r, r->t_min, max_dist, &t_enter, &t_exit)) return 1.0;
    double t = fmax(t_enter, r->t_min);
    const double limit = fmin(t_exit, max_dist);
    if (t >= limit) return 1.0;
    double tr = 1.0;
    const double inv_maj = 1.0 / m->majorant;
    for (int it = 0; it < 256; ++it) {
        const double u = fmax(1.0 - orng_next_f64(rng), 1e-15);
        t -= log(u) * inv_maj;
        if (t >= limit) break;
        const double p[3] = {r->o[0] + r->d[0] * t, r->o[1] + r->d[1] * t,
                             r->o[2] + r->d[2] * t};
        double rho, temp;
        medium_at(s, m, p, &rho, &temp);
        const double sigma_t = fmin(rho * m->sigma_t_scale, m->majorant);
        tr *= 1.0 - sigma_t * inv_maj;
        if (tr < 0.05) {
            const double q = fmin(1.0 - tr, 0.95);
            if (orng_next_f64(rng) < q) return 0.0;
            tr /= 1.0 - q;
        }
    }
    return rclamp(tr, 0.0, 1.0);
}

/* fluid.rs:878-888 phase_hg_pdf */
double of_phase_hg_pdf(double g, const double *wo, const double *wi) {
    if (fabs(g) < 1e-4) return 1.0 / (4.0 * O_PI);
    const double nwo[3] = {-wo[0], -wo[1], -wo[2]};
    const double cos_theta = rclamp(vdot(nwo, wi), -1.0, 1.0);
    const double denom = fmax(1.0 + g * g - 2.0 * g * cos_theta, 1e-12);
    return (1.0 - g * g) / (4.0 * O_PI * denom * sqrt(denom));
}

/* fluid.rs:892-910 sample_phase_hg; wi_pdf: wi x3, pdf */
void of_sample_phase_hg(double g, const double *wo, ORng *rng, double *wi_pdf) {
    if (fabs(g) < 1e-4) { /* rng.rs:133-138 */
        const double z = 1.0 - 2.0 * orng_next_f64(rng);
        const double rr = sqrt(fmax(1.0 - z * z, 0.0));
        const double phi = O_TAU * orng_next_f64(rng);
        wi_pdf[0] = rr * cos(phi);
        wi_pdf[1] = rr * sin(phi);
        wi_pdf[2] = z;
        wi_pdf[3] = 1.0 / (4.0 * O_PI);
        return;
    }
    const double sv = 2.0 * orng_next_f64(rng) - 1.0;
    const double sqr = (1.0 - g * g) / (1.0 + g * sv);
    const double cos_theta = rclamp((1.0 + g * g - sqr * sqr) / (2.0 * g), -1.0, 1.0);
    const double sin_theta = sqrt(fmax(1.0 - cos_theta * cos_theta, 0.0));
    const double phi = O_TAU * orng_next_f64(rng);
    const double nwo[3] = {-wo[0], -wo[1], -wo[2]};
    double fwd[3], w[3];
    vnormalize(nwo, fwd);
    vnormalize(fwd, w); /* Onb::from_normal normalises again (math.rs:728) */
    const double sign = copysign(1.0, w[2]);
    const double a = -1.0 / (sign + w[2]);
    const double b = w[0] * w[1] * a;
    const double u[3] = {1.0 + sign * w[0] * w[0] * a, sign * b, -sign * w[0]};
    const double v[3] = {b, sign + w[1] * w[1] * a, -w[1]};
    const double lx = sin_theta * cos(phi), ly = sin_theta * sin(phi), lz = cos_theta;
    const double world[3] = {u[0] * lx + v[0] * ly + w[0] * lz, u[1] * lx + v[1] * ly + w[1] * lz,
                             u[2] * lx + v[2] * ly + w[2] * lz};
    vnormalize(world, wi_pdf);
    wi_pdf[3] = of_phase_hg_pdf(g, wo, wi_pdf);
}

/* ------------------------------------------------------------------ */
/* Telemetry -> soundtrack (audio.rs:1-639), restated to pin the whole  */
/* step() + compute_telemetry chain against the reference's rendered    */
/* renders/simulation_soundtrack.wav (main.rs:82-94, 166-168).          */
/* ------------------------------------------------------------------ */
#define AU_SR 48000
#define AU_FRAMES (AU_SR * 15)
#define AU_TAU TAU_                                      /* std::f64::consts::TAU */
#define AU_PI 3.14159265358979323846264338327950288       /* std::f64::consts::PI */
#define AU_FRAC_1_SQRT_2 0.707106781186547524400844362104849039

typedef struct { double ic1, ic2; } AuSvf2;
typedef struct { AuSvf2 s1, s2; } AuSvf4;

/* audio.rs:27-47 TptSvf2Pole::process */
static void au_svf2(AuSvf2 *s, double in, double cutoff_hz, double res, double fs, double *lp,
                    double *bp) {
    double fc = rclamp(cutoff_hz, 20.0, 0.47 * fs);
    double g = tan(AU_PI * fc / fs);
    double k = 2.0 * (1.0 - rclamp(res, 0.0, 0.97));
    double a1 = 1.0 / (1.0 + g * (g + k));
    double a2 = g * a1;
    double a3 = g * a2;
    double v3 = in - s->ic2;
    double v1 = a1 * s->ic1 + a2 * v3;
    double v2 = s->ic2 + a2 * s->ic1 + a3 * v3;
    s->ic1 = 2.0 * v1 - s->ic1;
    s->ic2 = 2.0 * v2 - s->ic2;
    *lp = v2;
    *bp = v1;
}

/* audio.rs:59-65 */
static double au_lowpass4(AuSvf4 *s, double in, double fc, double res, double fs) {
    double lp1, bp1, lp2, bp2;
    au_svf2(&s->s1, in, fc, res * 0.88, fs, &lp1, &bp1);
    double driven = tanh(lp1 * 1.08) * (1.0 / 1.08);
    au_svf2(&s->s2, driven, fc, res, fs, &lp2, &bp2);
    return lp2;
}

/* audio.rs:67-72 */
static double au_bandpass4(AuSvf4 *s, double in, double fc, double res, double fs) {
    double lp1, bp1, lp2, bp2;
    au_svf2(&s->s1, in, fc, res * 0.85, fs, &lp1, &bp1);
    au_svf2(&s->s2, bp1, fc, res, fs, &lp2, &bp2);
    return bp2;
}

/* audio.rs:82-111 */
static void au_hadamard8(double *x) {
    for (int i = 0; i < 8; i += 2) {
        double a = x[i], b = x[i + 1];
        x[i] = a + b;
        x[i + 1] = a - b;
    }
    for (int base = 0; base < 8; base += 4)
        for (int j = 0; j < 2; ++j) {
            double a = x[base + j], b = x[base + j + 2];
            x[base + j] = a + b;
            x[base + j + 2] = a - b;
        }
    for (int j = 0; j < 4; ++j) {
        double a = x[j], b = x[j + 4];
        x[j] = a + b;
        x[j + 4] = a - b;
    }
    const double inv_sqrt8 = 0.5 * AU_FRAC_1_SQRT_2;
    for (int i = 0; i < 8; ++i) x[i] *= inv_sqrt8;
}

typedef struct {
    double *buf[8];
    size_t len[8], base[8], wp[8];
    double damp_state[8], gain_dc[8], damp_coeff[8], lfo_phase[8], lfo_inc[8];
} AuFdn;

/* audio.rs:125-159 HadamardFdn8::new */
static void au_fdn_init(AuFdn *f, double fs, double rt60_low, double rt60_high) {
    static const size_t base[8] = {1499, 1877, 2267, 2683, 3119, 3583, 4079, 4637};
    for (int i = 0; i < 8; ++i) {
        double d_sec = (double)base[i] / fs;
        double g_low = pow(10.0, -3.0 * d_sec / fmax(rt60_low, 0.2));
        double g_high = pow(10.0, -3.0 * d_sec / fmax(rt60_high, 0.1));
        f->gain_dc[i] = fmin(g_low, 0.992);
        f->damp_coeff[i] = rclamp(1.0 - (g_high / fmax(g_low, 1e-6)), 0.08, 0.72);
        f->lfo_phase[i] = (double)i * (AU_TAU / 8.0);
        f->lfo_inc[i] = AU_TAU * (0.21 + 0.07 * (double)i) / fs;
        f->base[i] = base[i];
        f->len[i] = base[i] + 64;
        f->buf[i] = (double *)calloc(f->len[i], sizeof(double));
        f->wp[i] = 0;
        f->damp_state[i] = 0.0;
    }
}

static void au_fdn_free(AuFdn *f) {
    for (int i = 0; i < 8; ++i) free(f->buf[i]);
}

/* audio.rs:161-215 HadamardFdn8::process_stereo */
static void au_fdn_process(AuFdn *f, double in_l, double in_r, double *wet_l, double *wet_r) {
    double delayed[8];
    for (int i = 0; i < 8; ++i) {
        const double *buf = f->buf[i];
        size_t len = f->len[i];
        double mod_samples = 6.5 * sin(f->lfo_phase[i]);
        f->lfo_phase[i] += f->lfo_inc[i];
        if (f->lfo_phase[i] >= AU_TAU) f->lfo_phase[i] -= AU_TAU;
        double read_offset = (double)f->base[i] + mod_samples;
        double read_f = (double)f->wp[i] + (double)len - read_offset;
        size_t r0 = ((size_t)floor(read_f)) % len;
        size_t r1 = (r0 + 1) % len;
        double frac = read_f - floor(read_f);
        double raw = buf[r0] * (1.0 - frac) + buf[r1] * frac;
        double lp = raw * (1.0 - f->damp_coeff[i]) + f->damp_state[i] * f->damp_coeff[i];
        f->damp_state[i] = lp;
        delayed[i] = lp * f->gain_dc[i];
    }
    double mixed[8];
    memcpy(mixed, delayed, sizeof(mixed));
    au_hadamard8(mixed);
    double inputs[8] = {in_l,          in_r,        0.7 * (in_l + in_r), 0.7 * (in_l - in_r),
                        -in_l * 0.8,   in_r * 0.8,  -0.6 * (in_l - in_r), 0.6 * (in_l + in_r)};
    for (int i = 0; i < 8; ++i) {
        size_t wp = f->wp[i];
        f->buf[i][wp] = mixed[i] + inputs[i] * 0.35;
        f->wp[i] = (wp + 1) % f->len[i];
    }
    *wet_l = 0.38 * (delayed[0] - delayed[2] + delayed[4] - delayed[6]);
    *wet_r = 0.38 * (delayed[1] - delayed[3] + delayed[5] - delayed[7]);
}

typedef struct {
    int active;
    double freq_hz, age_sec, duration_sec, velocity, pan;
    double phase_c1, phase_m1, phase_c2, phase_m2, prev_m1;
} AuVoice;

/* audio.rs:241-257 (Default) */
static void au_voice_default(AuVoice *v) {
    memset(v, 0, sizeof(*v));
    v->freq_hz = 440.0;
    v->duration_sec = 3.5;
}

/* audio.rs:260-272 */
static void au_voice_trigger(AuVoice *v, double freq, double vel, double pan, double dur) {
    v->active = 1;
    v->freq_hz = freq;
    v->age_sec = 0.0;
    v->duration_sec = dur;
    v->velocity = vel;
    v->pan = rclamp(pan, -0.95, 0.95);
    v->phase_c1 = v->phase_m1 = v->phase_c2 = v->phase_m2 = v->prev_m1 = 0.0;
}

/* audio.rs:274-316 */
static void au_voice_render(AuVoice *v, double dt, double ke, double *l, double *r) {
    *l = *r = 0.0;
    if (!v->active) return;
    double t = v->age_sec;
    if (t >= v->duration_sec) {
        v->active = 0;
        return;
    }
    double attack = fmin(t / 0.0025, 1.0);
    double env_body = attack * exp(-t * 1.45);
    double env_mod1 = exp(-t * 4.8);
    double env_chime = attack * exp(-t * 3.1);
    double env_mod2 = exp(-t * 11.5);
    double tail_gate = rclamp((v->duration_sec - t) / 0.06, 0.0, 1.0);
    double idx1 = (1.35 + 2.45 * ke) * env_mod1;
    double idx2 = (0.85 + 2.10 * ke) * env_mod2;
    double m1 = sin(v->phase_m1 + 0.22 * v->prev_m1);
    v->prev_m1 = m1;
    double c1 = sin(v->phase_c1 + idx1 * m1) * env_body;
    double m2 = sin(v->phase_m2);
    double c2 = sin(v->phase_c2 + idx2 * m2) * env_chime * 0.36;
    v->phase_c1 = fmod(v->phase_c1 + AU_TAU * v->freq_hz * dt, AU_TAU);
    v->phase_m1 = fmod(v->phase_m1 + AU_TAU * (v->freq_hz * 2.0007) * dt, AU_TAU);
    v->phase_c2 = fmod(v->phase_c2 + AU_TAU * (v->freq_hz * 2.0015) * dt, AU_TAU);
    v->phase_m2 = fmod(v->phase_m2 + AU_TAU * (v->freq_hz * 5.4142) * dt, AU_TAU);
    v->age_sec += dt;
    double mono = (c1 + c2) * v->velocity * tail_gate;
    double angle = (v->pan + 1.0) * (AU_PI * 0.25);
    *l = mono * cos(angle);
    *r = mono * sin(angle);
}

typedef struct { double buf[128]; size_t wp; } AuMicroDelay;

/* audio.rs:319-338 */
static double au_micro_delay(AuMicroDelay *m, double x, double delay_samples) {
    m->buf[m->wp] = x;
    double d = rclamp(delay_samples, 0.0, 64.0);
    double rf = (double)m->wp + 128.0 - d;
    size_t r0 = ((size_t)floor(rf)) & 127;
    size_t r1 = (r0 + 1) & 127;
    double frac = rf - floor(rf);
    m->wp = (m->wp + 1) & 127;
    return m->buf[r0] * (1.0 - frac) + m->buf[r1] * frac;
}

static double au_midi_to_hz(double midi) { return 440.0 * pow(2.0, (midi - 69.0) / 12.0); }

/* audio.rs:344-355 */
static double au_poly_blep(double t, double dt) {
    if (t < dt) {
        double x = t / dt;
        return x + x - x * x - 1.0;
    } else if (t > 1.0 - dt) {
        double x = (t - 1.0) / dt;
        return x * x + x + x + 1.0;
    }
    return 0.0;
}

static double au_fract(double x) { return x - trunc(x); }

/* audio.rs:358-394; frames hold 8 values per step (of_telemetry order) */
static void au_curves(const double *fr, long n, double t_sec, double max_ke, double max_vort,
                      double *ke_n, double *vort_n, double *cx_o, double *cy_o) {
    if (n == 1) {
        *ke_n = rclamp(fr[1] / fmax(max_ke, 1e-9), 0.0, 1.0);
        *vort_n = rclamp(fr[3] / fmax(max_vort, 1e-9), 0.0, 1.0);
        *cx_o = rclamp(fr[5], -1.0, 1.0);
        *cy_o = rclamp(0.5 * (fr[6] + 1.0), 0.0, 1.0);
        return;
    }
    double u = rclamp(t_sec / 15.0, 0.0, 1.0) * (double)(n - 1);
    long i0 = (long)floor(u);
    if (i0 > n - 1) i0 = n - 1;
    long i1 = i0 + 1 < n - 1 ? i0 + 1 : n - 1;
    double frac = u - (double)i0;
    const double *a = fr + 8 * i0, *b = fr + 8 * i1;
    double ke = a[1] * (1.0 - frac) + b[1] * frac;
    double vort = a[3] * (1.0 - frac) + b[3] * frac;
    double cx = a[5] * (1.0 - frac) + b[5] * frac;
    double cy = a[6] * (1.0 - frac) + b[6] * frac;
    *ke_n = rclamp(ke / fmax(max_ke, 1e-9), 0.0, 1.0);
    *vort_n = rclamp(vort / fmax(max_vort, 1e-9), 0.0, 1.0);
    *cx_o = rclamp(cx, -1.0, 1.0);
    *cy_o = rclamp(0.5 * (cy + 1.0), 0.0, 1.0);
}

/* audio.rs:397-560 synthesize_fluid_soundtrack, then the TPDF-dithered
 * 24-bit quantisation of audio.rs:593-604: pcm receives 2 * nsamples values
 * (left, right interleaved; nsamples <= 720000).  max_ke / max_vort <= 0:
 * the normalisers of audio.rs:401-406 (the maxima over the frames); > 0:
 * used as given. */
void of_soundtrack_ex(const double *frames, long n, double max_ke_in, double max_vort_in,
                      long nsamples, int32_t *pcm) {
    const double fs = (double)AU_SR, dt = 1.0 / fs;
    double max_ke = 1e-6, max_vort = 1e-6;
    for (long a = 0; a < n; ++a) {
        max_ke = fmax(max_ke, frames[8 * a + 1]);
        max_vort = fmax(max_vort, frames[8 * a + 3]);
    }
    if (max_ke_in > 0.0) max_ke = max_ke_in;
    if (max_vort_in > 0.0) max_vort = max_vort_in;
    if (nsamples > AU_FRAMES) nsamples = AU_FRAMES;
    AuVoice voices[16];
    for (int v = 0; v < 16; ++v) au_voice_default(&voices[v]);
    AuSvf4 pad_l_f = {{0, 0}, {0, 0}}, pad_r_f = {{0, 0}, {0, 0}};
    AuSvf4 br_l_f = {{0, 0}, {0, 0}}, br_r_f = {{0, 0}, {0, 0}};
    AuFdn fdn;
    au_fdn_init(&fdn, fs, 3.85, 1.35);
    AuMicroDelay itd_l, itd_r;
    memset(&itd_l, 0, sizeof(itd_l));
    memset(&itd_r, 0, sizeof(itd_r));
    ORng rng;
    orng_seed(&rng, 0xAE74E9C02026ull);
    static const double chord_pads[4][4] = {
        {38.0, 57.0, 60.0, 64.0}, {34.0, 53.0, 57.0, 64.0}, {31.0, 50.0, 58.0, 62.0},
        {33.0, 52.0, 57.0, 64.0}};
    static const double bell_scales[4][8] = {
        {62.0, 64.0, 65.0, 69.0, 72.0, 76.0, 79.0, 81.0},
        {58.0, 62.0, 64.0, 65.0, 69.0, 72.0, 76.0, 77.0},
        {55.0, 58.0, 62.0, 65.0, 69.0, 70.0, 74.0, 79.0},
        {57.0, 62.0, 64.0, 67.0, 69.0, 74.0, 76.0, 81.0}};
    static const size_t pattern[8] = {0, 3, 4, 2, 6, 4, 5, 7};
    double pad_phases[4][3] = {{0}};
    const double detunes[3] = {0.9955, 1.0, 1.0048};
    double sub_phase = 0.0;
    double smooth_ke = 0.25, smooth_vort = 0.25, prev_ke = 0.25;
    size_t next_grid_step = 0, step_counter = 0;
    double flux_cooldown = 0.0;
    const size_t step_interval = (size_t)AU_SR * 3 / 16;
    ORng dither;
    orng_seed(&dither, 0x24B17DA7A202ull);
    for (size_t fi = 0; fi < (size_t)nsamples; ++fi) {
        double t_sec = (double)fi * dt;
        size_t chord_idx = (size_t)floor(t_sec / 3.75);
        if (chord_idx > 3) chord_idx = 3;
        double raw_ke, raw_vort, com_x, com_y;
        au_curves(frames, n, t_sec, max_ke, max_vort, &raw_ke, &raw_vort, &com_x, &com_y);
        smooth_ke += 0.0015 * (raw_ke - smooth_ke);
        smooth_vort += 0.0015 * (raw_vort - smooth_vort);
        double ke_flux = (smooth_ke - prev_ke) * fs;
        prev_ke = smooth_ke;
        if (flux_cooldown > 0.0) flux_cooldown -= dt;
        int grid_hit = fi >= next_grid_step && t_sec < 14.15;
        int flux_hit = ke_flux > 0.22 && flux_cooldown <= 0.0 && t_sec > 0.4 && t_sec < 14.0;
        if (grid_hit || flux_hit) {
            if (grid_hit) next_grid_step += step_interval;
            if (flux_hit) flux_cooldown = 0.14;
            const double *scale = bell_scales[chord_idx];
            size_t pattern_offset = pattern[step_counter & 7];
            size_t fluid_offset = (size_t)round(com_y * 3.0);
            if (fluid_offset > 3) fluid_offset = 3;
            size_t degree = (pattern_offset + fluid_offset) & 7;
            double midi = scale[degree];
            double vel = fmin(0.16 + 0.18 * smooth_ke + (flux_hit ? 0.06 : 0.0), 0.36);
            double spread = (step_counter & 1) == 0 ? -0.45 : 0.45;
            double pan = rclamp(0.65 * com_x * 3.0 + spread, -0.92, 0.92);
            size_t slot = 0;
            double best_age = -1.0;
            for (size_t vi = 0; vi < 16; ++vi) {
                if (!voices[vi].active) {
                    slot = vi;
                    break;
                }
                if (voices[vi].age_sec > best_age) {
                    best_age = voices[vi].age_sec;
                    slot = vi;
                }
            }
            au_voice_trigger(&voices[slot], au_midi_to_hz(midi), vel, pan, 3.1 + 1.2 * smooth_ke);
            step_counter += 1;
        }
        double bells_l = 0.0, bells_r = 0.0;
        for (int v = 0; v < 16; ++v) {
            double vl, vr;
            au_voice_render(&voices[v], dt, smooth_ke, &vl, &vr);
            bells_l += vl;
            bells_r += vr;
        }
        const double *chord = chord_pads[chord_idx];
        double pad_raw_l = 0.0, pad_raw_r = 0.0;
        for (int ni = 0; ni < 4; ++ni) {
            double base_f = au_midi_to_hz(chord[ni]);
            for (int di = 0; di < 3; ++di) {
                double f = base_f * detunes[di];
                double p_inc = f * dt;
                double ph = pad_phases[ni][di];
                double saw = 2.0 * ph - 1.0 - au_poly_blep(ph, p_inc);
                pad_phases[ni][di] = au_fract(ph + p_inc);
                if (di == 0) {
                    pad_raw_l += saw * 0.095;
                    pad_raw_r += saw * 0.035;
                } else if (di == 2) {
                    pad_raw_l += saw * 0.035;
                    pad_raw_r += saw * 0.095;
                } else {
                    pad_raw_l += saw * 0.065;
                    pad_raw_r += saw * 0.065;
                }
            }
        }
        double sub_f = au_midi_to_hz(chord[0]);
        double sub_sig = sin(AU_TAU * sub_phase) * (0.19 + 0.08 * smooth_ke);
        sub_phase = au_fract(sub_phase + sub_f * dt);
        double lfo_breath = 0.5 + 0.5 * sin(AU_TAU * 0.18 * t_sec);
        double cutoff_hz = 185.0 * pow(2.0, 3.9 * smooth_ke + 0.55 * lfo_breath);
        double resonance = rclamp(0.24 + 0.48 * smooth_vort, 0.15, 0.78);
        double pad_l = au_lowpass4(&pad_l_f, pad_raw_l, cutoff_hz, resonance, fs) + sub_sig;
        double pad_r = au_lowpass4(&pad_r_f, pad_raw_r, cutoff_hz * 1.03, resonance, fs) + sub_sig;
        double white_l = 2.0 * orng_next_f64(&rng) - 1.0;
        double white_r = 2.0 * orng_next_f64(&rng) - 1.0;
        double air_fc = 620.0 + 2400.0 * com_y + 1200.0 * smooth_vort;
        double air_amp = 0.022 * smooth_vort;
        double air_l = au_bandpass4(&br_l_f, white_l, air_fc, 0.68, fs) * air_amp;
        double air_r = au_bandpass4(&br_r_f, white_r, air_fc * 1.06, 0.68, fs) * air_amp;
        double azimuth = rclamp(com_x * 2.5, -1.0, 1.0);
        double delay_l = azimuth > 0.0 ? azimuth * 24.0 : 0.0;
        double delay_r = azimuth < 0.0 ? -azimuth * 24.0 : 0.0;
        double dry_l = au_micro_delay(&itd_l, bells_l + air_l, delay_l) + pad_l * 0.78;
        double dry_r = au_micro_delay(&itd_r, bells_r + air_r, delay_r) + pad_r * 0.78;
        double rev_l = bells_l * 0.72 + pad_l * 0.42 + air_l * 0.85;
        double rev_r = bells_r * 0.72 + pad_r * 0.42 + air_r * 0.85;
        double wet_l, wet_r;
        au_fdn_process(&fdn, rev_l, rev_r, &wet_l, &wet_r);
        double fade_in = rclamp(t_sec / 0.08, 0.0, 1.0);
        double fade_out = 1.0;
        if (t_sec > 14.35) {
            double u = rclamp((15.0 - t_sec) / 0.65, 0.0, 1.0);
            fade_out = 0.5 - 0.5 * cos(AU_PI * u);
        }
        double master = fade_in * fade_out;
        double mix[2];
        mix[0] = tanh((dry_l * 0.74 + wet_l * 0.52) * master * 0.92) * 0.91;
        mix[1] = tanh((dry_r * 0.74 + wet_r * 0.52) * master * 0.92) * 0.91;
        for (int c = 0; c < 2; ++c) { /* audio.rs:593-604 */
            double a = orng_next_f64(&dither);
            double b = orng_next_f64(&dither);
            double tpdf = (a - b) * 0.5;
            double v = rclamp(round(rclamp(mix[c], -1.0, 1.0) * 8388607.0 + tpdf), -8388608.0,
                              8388607.0);
            pcm[2 * fi + c] = (int32_t)v;
        }
    }
    au_fdn_free(&fdn);
}

void of_soundtrack(const double *frames, long n, int32_t *pcm) {
    of_soundtrack_ex(frames, n, 0.0, 0.0, AU_FRAMES, pcm);
}

/* Batched forms of the volumetric samplers above, for the parity tests: ray
 * a has origin o[3a..], direction d[3a..] (Ray::dir, already normalised;
 * inv_dir = 1/dir as Ray::new forms it), t_min[a], far limit t_far[a], and
 * its own RNG state states[4a..4a+3], advanced in place. */
static void batch_ray(const double *o, const double *d, const double *t_min, long a, ORay *r) {
    for (int q = 0; q < 3; ++q) {
        r->o[q] = o[3 * a + q];
        r->d[q] = d[3 * a + q];
        r->inv[q] = 1.0 / d[3 * a + q];
    }
    r->t_min = t_min[a];
}

void of_medium_interaction_batch(const OFluid *s, const OMedium *m, long n, const double *o,
                                 const double *d, const double *t_min, const double *t_far,
                                 const double *lambda, uint64_t *states, double *out) {
    for (long a = 0; a < n; ++a) {
        ORay r;
        ORng g;
        batch_ray(o, d, t_min, a, &r);
        memcpy(g.s, states + 4 * a, sizeof(g.s));
        of_medium_sample_interaction(s, m, &r, t_far[a], lambda + 4 * a, &g, out + 14 * a);
        memcpy(states + 4 * a, g.s, sizeof(g.s));
    }
}

void of_medium_transmittance_batch(const OFluid *s, const OMedium *m, long n, const double *o,
                                   const double *d, const double *t_min, const double *t_far,
                                   uint64_t *states, double *out) {
    for (long a = 0; a < n; ++a) {
        ORay r;
        ORng g;
        batch_ray(o, d, t_min, a, &r);
        memcpy(g.s, states + 4 * a, sizeof(g.s));
        out[a] = of_medium_transmittance(s, m, &r, t_far[a], &g);
        memcpy(states + 4 * a, g.s, sizeof(g.s));
    }
}

void of_phase_sample_batch(double g, long n, const double *wo, uint64_t *states, double *wi,
                           double *pdf) {
    for (long a = 0; a < n; ++a) {
        ORng r;
        double out[4];
        memcpy(r.s, states + 4 * a, sizeof(r.s));
        of_sample_phase_hg(g, wo + 3 * a, &r, out);
        memcpy(states + 4 * a, r.s, sizeof(r.s));
        wi[3 * a] = out[0];
        wi[3 * a + 1] = out[1];
        wi[3 * a + 2] = out[2];
        pdf[a] = out[3];
    }
}

void of_phase_pdf_batch(double g, long n, const double *wo, const double *wi, double *pdf) {
    for (long a = 0; a < n; ++a) pdf[a] = of_phase_hg_pdf(g, wo + 3 * a, wi + 3 * a);
}

/* Vec3::normalize (math.rs:58-65) of n vectors, as Ray::new applies to dir */
void of_normalize_batch(long n, const double *v, double *out) {
    for (long a = 0; a < n; ++a) vnormalize(v + 3 * a, out + 3 * a);
}
