/*
 * sanitize_check.c -- test infrastructure only: drives the f64 oracle
 * (fluid_oracle.c) through every public function under AddressSanitizer and
 * UndefinedBehaviorSanitizer (built by `make -C oracle sanitize`, run by
 * tests/test_oracle.py).  Any out-of-bounds access, leak or undefined
 * behaviour in the checker aborts with a non-zero exit status.  The
 * scenarios are small versions of the ones the parity tests use: the
 * src/verify.rs:229-250 projection self-check, reference-default steps with
 * the jets, sphere-source steps with CG and red-black SOR, all five
 * advection kinds on a random velocity, point and medium sampling inside and
 * outside the clamp box, and the telemetry-driven soundtrack.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct OFluid OFluid;
typedef struct {
    uint64_t s[4];
} ORng;

OFluid *of_create(long nx, long ny, long nz, double h_override);
void of_destroy(OFluid *s);
void of_set_params(OFluid *s, double eps, double beta, double alpha, int solver_kind,
                   long solver_iters, double omega, double rel_tol, double diss_vel,
                   double diss_rho, double diss_temp);
void of_set_source(OFluid *s, int kind, double cx, double cy, double cz, double r, double rho,
                   double temp);
double *of_field(OFluid *s, int id);
double of_sample_trilinear(const OFluid *s, int field, double x, double y, double z);
double of_max_density(const OFluid *s);
void of_medium_sample(const OFluid *s, const double *bmin, const double *bmax, double px,
                      double py, double pz, double *d_out, double *t_out);
void of_advect(OFluid *s, int kind, double dt, double dissipation);
void of_enforce_bc(OFluid *s);
void of_buoyancy_confinement(OFluid *s, double dt);
double of_divergence_norm(OFluid *s);
double of_divergence_max(const OFluid *s);
void of_apply_laplacian(const OFluid *s, const double *x, double *out);
void of_project(OFluid *s);
void of_inject_sources(OFluid *s, double time_sec, double dt);
void of_telemetry(const OFluid *s, double time_sec, double *out);
void of_step(OFluid *s, double dt, double time_sec, double *telemetry_out);
void orng_seed(ORng *r, uint64_t seed);
double orng_range_f64(ORng *r, double lo, double hi);
void of_soundtrack_ex(const double *frames, long n, double max_ke_in, double max_vort_in,
                      long nsamples, int32_t *pcm);

static void defaults(OFluid *s, int kind, long iters) {
    /* src/fluid.rs:79-83, 398, 599-608 */
    of_set_params(s, 5.5, 6.8, 0.45, kind, iters, 1.65, 0.0, 0.996, 0.996, 0.985);
}

static int fail(const char *what) {
    fprintf(stderr, "sanitize_check: %s\n", what);
    return 1;
}

int main(void) {
    char buf[64];
    /* src/verify.rs:229-250: 2.5751e0 -> 6.7289e-3 (README.md:54) */
    OFluid *s = of_create(16, 20, 16, 0.0);
    defaults(s, 0, 40);
    of_inject_sources(s, 0.5, 0.1);
    of_buoyancy_confinement(s, 0.1);
    const double pre = of_divergence_norm(s);
    of_project(s);
    const double post = of_divergence_norm(s);
    snprintf(buf, sizeof buf, "%.4e %.4e", pre, post);
    printf("known answer %s\n", buf);
    if (strcmp(buf, "2.5751e+00 6.7289e-03") != 0) return fail("known answer mismatch");

    /* reference-default steps with the jets, and the telemetry frames */
    double frames[8 * 9];
    for (int t = 0; t < 9; ++t) of_step(s, 0.08, 15.0 * t / 119.0, frames + 8 * t);
    if (!(frames[8 * 8 + 1] > 0.0)) return fail("kinetic energy not positive");
    int32_t *pcm = malloc(sizeof(int32_t) * 2 * 4800);
    if (!pcm) return fail("out of memory");
    of_soundtrack_ex(frames, 9, 0.0, 0.0, 4800, pcm);
    free(pcm);
    of_destroy(s);

    /* sphere source, CG to a relative tolerance and red-black SOR */
    for (int kind = 0; kind < 2; ++kind) {
        OFluid *q = of_create(32, 48, 32, 1.0 / 32.0);
        of_set_params(q, 0.3, 1.0, 0.05, kind, kind ? 40 : 200, 1.7, 1e-5, 0.996, 0.996,
                      0.985);
        of_set_source(q, 1, 16.0, 4.0, 16.0, 4.0, 1.0, 1.0);
        double tel[8];
        for (int t = 0; t < 3; ++t) of_step(q, 0.02, 0.02 * t, tel);
        if (!(of_divergence_max(q) >= 0.0)) return fail("divergence max");
        of_destroy(q);
    }

    /* every advection kind on a random state, the Laplacian, sampling */
    const long nx = 13, ny = 11, nz = 9, n = nx * ny * nz;
    OFluid *r = of_create(nx, ny, nz, 0.0);
    defaults(r, 0, 32);
    ORng rng;
    orng_seed(&rng, 7);
    for (int f = 0; f < 6; ++f) {
        double *a = of_field(r, f);
        for (long c = 0; c < n; ++c) a[c] = orng_range_f64(&rng, f < 3 ? -3.0 : 0.0, 3.0);
    }
    for (int kind = 0; kind < 5; ++kind) of_advect(r, kind, 0.5, kind == 4 ? 0.985 : 0.996);
    of_enforce_bc(r);
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    if (!x || !y) return fail("out of memory");
    for (long c = 0; c < n; ++c) x[c] = orng_range_f64(&rng, -1.0, 1.0);
    of_apply_laplacian(r, x, y);
    free(x);
    free(y);
    const double pts[5][3] = {{-4.0, 2.0, 3.0}, {0.5, 0.5, 0.5}, {6.2, 5.1, 4.9},
                              {12.0, 10.0, 8.0}, {30.0, -7.0, 1e6}};
    const double bmin[3] = {-1.0, -1.0, -1.0}, bmax[3] = {1.0, 1.0, 1.0};
    for (int p = 0; p < 5; ++p) {
        double d = 0.0, t = 0.0;
        (void)of_sample_trilinear(r, 3, pts[p][0], pts[p][1], pts[p][2]);
        of_medium_sample(r, bmin, bmax, pts[p][0] / 10.0, pts[p][1] / 10.0, pts[p][2] / 10.0, &d,
                         &t);
    }
    double tel[8];
    of_telemetry(r, 1.0, tel);
    (void)of_max_density(r);
    of_destroy(r);
    printf("sanitize_check: ok\n");
    return 0;
}
