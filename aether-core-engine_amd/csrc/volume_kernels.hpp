// volume_kernels.hpp -- the renderer-facing sampling of FluidMedium
// (src/fluid.rs:739-910) on the device, in FP64, one thread per ray.
//
// Each ray carries its own Xoshiro256++ state (src/rng.rs:48-77), loaded at
// the start and stored back advanced, so a batch of rays consumes exactly
// the draws the reference's per-ray `&mut Rng` calls would.  Expressions
// keep the reference's association, including Vec3::dot's fused
// multiply-adds (src/math.rs:34-36); log/exp/sin/cos are the device's.
#pragma once
#include <stdint.h>

#include "aether_fluid.h"
#include "fluid_kernels.hpp"

namespace aether {
namespace {

// Medium parameters fixed at FluidMedium::from_solver (src/fluid.rs:703-735).
struct MediumDev {
    Dims g;
    Box3 box;
    double sigma_t_scale;
    double max_density;
    double majorant;  // max(max_density * sigma_t_scale, 1e-4)
    double phase_g;   // clamped to [-0.9, 0.9]
};

// One ray as Ray::new builds it (src/math.rs:763-773), plus the far limit
// of the query (surface_t of sample_interaction, max_dist of transmittance).
typedef aether_medium_ray RayIn;

__device__ __forceinline__ uint64_t xo_rotl(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
}

// Xoshiro256++ (src/rng.rs:48-64) and next_f64 (rng.rs:74-77).
struct DevRng {
    uint64_t s[4];
    __device__ __forceinline__ uint64_t next_u64() {
        const uint64_t result = xo_rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = xo_rotl(s[3], 45);
        return result;
    }
    __device__ __forceinline__ double next_f64() {
        return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0);
    }
};

__device__ __forceinline__ double vdot3(const double *a, const double *b) {  // math.rs:34-36
    return fma(a[0], b[0], fma(a[1], b[1], a[2] * b[2]));
}

__device__ __forceinline__ void vnormalize3(const double *v, double *out) {  // math.rs:58-65
    const double len = sqrt(vdot3(v, v));
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

// Aabb::intersect (math.rs:859-893); f64::min/max skip a NaN operand, as
// fmin/fmax do (an axis-parallel ray has an infinite inv_dir).
__device__ __forceinline__ bool aabb_hit(const Box3 &b, const double *o, const double *inv,
                                         double t_min, double t_max, double &t0, double &t1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double ta = (b.lo[a] - o[a]) * inv[a];
        const double tb = (b.hi[a] - o[a]) * inv[a];
        t_min = fmax(t_min, fmin(ta, tb));
        t_max = fmin(t_max, fmax(ta, tb));
        if (a < 2 && t_max < t_min) return false;
    }
    t0 = t_min;
    t1 = t_max;
    return t_max >= t_min;
}

// SpectralCurve::eval_at (math.rs:1148-1170) for FluidMedium's two curves
// (fluid.rs:721-731): neon = 0 is the SmoothRgb scatter albedo, 1 the
// NeonDischarge hot-core emission.
__device__ __forceinline__ double spectral_at(int neon, double lambda) {
    const double w_b = 1.0 / (1.0 + exp((lambda - 492.0) / 14.0));
    const double w_r = 1.0 / (1.0 + exp(-(lambda - 586.0) / 14.0));
    const double w_g = fmax(1.0 - w_b - w_r, 0.0);
    if (!neon) return 0.84 * w_r + 0.90 * w_g + 0.96 * w_b;
    const double base = 0.18 * w_r + 0.72 * w_g + 1.15 * w_b;
    const double e1 = (lambda - 486.0) / 14.0, e2 = (lambda - 618.0) / 16.0;
    const double g1 = 1.6 * exp(-0.5 * (e1 * e1));
    const double g2 = 1.1 * exp(-0.5 * (e2 * e2));
    return base + g1 + g2;
}

// sample_density_and_temp (fluid.rs:739-788) at p.
template <typename T>
__device__ __forceinline__ void medium_at(const MediumDev &m, const T *__restrict__ rho,
                                          const T *__restrict__ temp, const double *p,
                                          double &d_out, double &t_out) {
    const Box3 &box = m.box;
    if (p[0] < box.lo[0] || p[0] > box.hi[0] || p[1] < box.lo[1] || p[1] > box.hi[1] ||
        p[2] < box.lo[2] || p[2] > box.hi[2]) {
        d_out = 0.0;
        t_out = 0.0;
        return;
    }
    double u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ext = fmax(box.hi[d] - box.lo[d], 0.0);
        u[d] = (p[d] - box.lo[d]) / fmax(ext, 1e-9);
    }
    const double fx = rclamp_d(4.0 * u[0] * (1.0 - u[0]), 0.0, 1.0);
    const double fz = rclamp_d(4.0 * u[2] * (1.0 - u[2]), 0.0, 1.0);
    const double fy = rclamp_d(1.0 - fmax(u[1] - 0.95, 0.0) / 0.05, 0.0, 1.0);
    const double feather = sqrt(fx * fz * fy);
    const Dims &g = m.g;
    const int n[3] = {g.nx, g.ny, g.nz};
    int i0[3];
    double c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x = rclamp_d(u[d] * (double)(n[d] - 1), 0.5, (double)n[d] - 1.500001);
        const double fl = floor(x);
        i0[d] = min(max((int)fl, 0), n[d] - 2);  // NaN-safe index
        c[d] = x - fl;
    }
    const int b = (i0[2] * g.ny + i0[1]) * g.nx + i0[0];
    d_out = lerp3<double>(rho[b], rho[b + 1], rho[b + g.nx], rho[b + g.nx + 1], rho[b + g.sxy],
                          rho[b + g.sxy + 1], rho[b + g.sxy + g.nx], rho[b + g.sxy + g.nx + 1],
                          c[0], c[1], c[2]) *
            feather;
    t_out = lerp3<double>(temp[b], temp[b + 1], temp[b + g.nx], temp[b + g.nx + 1],
                          temp[b + g.sxy], temp[b + g.sxy + 1], temp[b + g.sxy + g.nx],
                          temp[b + g.sxy + g.nx + 1], c[0], c[1], c[2]) *
            feather;
}

__device__ __forceinline__ void load_ray(const RayIn &r, double *o, double *d, double *inv) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        o[a] = r.origin[a];
        d[a] = r.dir[a];
        inv[a] = 1.0 / d[a];  // Ray::new (math.rs:765)
    }
}

__device__ __forceinline__ DevRng load_rng(const uint64_t *st, long a) {
    DevRng r;
#pragma unroll
    for (int q = 0; q < 4; ++q) r.s[q] = st[4 * a + q];
    return r;
}
__device__ __forceinline__ void store_rng(const DevRng &r, uint64_t *st, long a) {
#pragma unroll
    for (int q = 0; q < 4; ++q) st[4 * a + q] = r.s[q];
}

// Woodcock delta tracking, fluid.rs:793-838: one MediumInteraction per ray
// (hit 0 for None).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_medium_interaction(
    MediumDev m, const T *__restrict__ rho, const T *__restrict__ temp,
    const RayIn *__restrict__ rays, const double *__restrict__ lambda, long nrays,
    uint64_t *__restrict__ rng_state, aether_medium_interaction *__restrict__ out) {
    const long a = (long)blockIdx.x * kThreads + tid_lin();
    if (a >= nrays) return;
    aether_medium_interaction r = {};
    out[a] = r;
    if (m.max_density <= 1e-6) return;
    const RayIn ray = rays[a];
    double o[3], d[3], inv[3];
    load_ray(ray, o, d, inv);
    double t_enter, t_exit;
    if (!aabb_hit(m.box, o, inv, ray.t_min, ray.t_far, t_enter, t_exit)) return;
    double t = fmax(t_enter, ray.t_min);
    const double limit = fmin(t_exit, ray.t_far);
    if (t >= limit) return;
    DevRng rng = load_rng(rng_state, a);
    const double inv_maj = 1.0 / m.majorant;
    for (int it = 0; it < 256; ++it) {
        const double u = fmax(1.0 - rng.next_f64(), 1e-15);
        t -= log(u) * inv_maj;
        if (t >= limit) break;
        const double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
        double dens, tmp;
        medium_at<T>(m, rho, temp, p, dens, tmp);
        const double sigma_t = dens * m.sigma_t_scale;
        if (rng.next_f64() * m.majorant < sigma_t) {
            r.hit = 1;
            r.t = t;
            r.point[0] = p[0];
            r.point[1] = p[1];
            r.point[2] = p[2];
            const double gb = fmax(tmp - 0.35, 0.0);
            const double glow = gb * gb * 0.95;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const double lam = lambda[4 * a + w];
                r.albedo[w] = spectral_at(0, lam);
                r.emission[w] = glow > 1e-5 ? spectral_at(1, lam) * glow : 0.0;
            }
            out[a] = r;
            break;
        }
    }
    store_rng(rng, rng_state, a);
}

// Ratio-tracking transmittance with Russian roulette, fluid.rs:841-874.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_medium_transmittance(
    MediumDev m, const T *__restrict__ rho, const T *__restrict__ temp,
    const RayIn *__restrict__ rays, long nrays, uint64_t *__restrict__ rng_state,
    double *__restrict__ out) {
    const long a = (long)blockIdx.x * kThreads + tid_lin();
    if (a >= nrays) return;
    out[a] = 1.0;
    if (m.max_density <= 1e-6) return;
    const RayIn ray = rays[a];
    double o[3], d[3], inv[3];
    load_ray(ray, o, d, inv);
    double t_enter, t_exit;
    if (!aabb_hit(m.box, o, inv, ray.t_min, ray.t_far, t_enter, t_exit)) return;
    double t = fmax(t_enter, ray.t_min);
    const double limit = fmin(t_exit, ray.t_far);
    if (t >= limit) return;
    DevRng rng = load_rng(rng_state, a);
    double tr = 1.0;
    bool killed = false;
    const double inv_maj = 1.0 / m.majorant;
    for (int it = 0; it < 256; ++it) {
        const double u = fmax(1.0 - rng.next_f64(), 1e-15);
        t -= log(u) * inv_maj;
        if (t >= limit) break;
        const double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
        double dens, tmp;
        medium_at<T>(m, rho, temp, p, dens, tmp);
        const double sigma_t = fmin(dens * m.sigma_t_scale, m.majorant);
        tr *= 1.0 - sigma_t * inv_maj;
        if (tr < 0.05) {
            const double q = fmin(1.0 - tr, 0.95);
            if (rng.next_f64() < q) {
                killed = true;
                break;
            }
            tr /= 1.0 - q;
        }
    }
    out[a] = killed ? 0.0 : rclamp_d(tr, 0.0, 1.0);
    store_rng(rng, rng_state, a);
}

// phase_hg_pdf, fluid.rs:878-888.
__device__ __forceinline__ double phase_hg(double g, const double *wo, const double *wi) {
    const double kFourPi = 4.0 * 3.14159265358979323846;
    if (fabs(g) < 1e-4) return 1.0 / kFourPi;
    const double nwo[3] = {-wo[0], -wo[1], -wo[2]};
    const double cos_theta = rclamp_d(vdot3(nwo, wi), -1.0, 1.0);
    const double denom = fmax(1.0 + g * g - 2.0 * g * cos_theta, 1e-12);
    return (1.0 - g * g) / (kFourPi * denom * sqrt(denom));
}

__global__ __launch_bounds__(kThreads) void k_phase_pdf(double g, const double *__restrict__ wo,
                                                        const double *__restrict__ wi, long n,
                                                        double *__restrict__ pdf) {
    const long a = (long)blockIdx.x * kThreads + tid_lin();
    if (a >= n) return;
    pdf[a] = phase_hg(g, wo + 3 * a, wi + 3 * a);
}

// sample_phase_hg, fluid.rs:892-910 (Rng::uniform_sphere, rng.rs:133-138,
// when |g| < 1e-4; Onb::from_normal / to_world, math.rs:727-740).
__global__ __launch_bounds__(kThreads) void k_phase_sample(double g, const double *__restrict__ wo,
                                                           long n, uint64_t *__restrict__ rng_state,
                                                           double *__restrict__ wi,
                                                           double *__restrict__ pdf) {
    const long a = (long)blockIdx.x * kThreads + tid_lin();
    if (a >= n) return;
    const double kTau = 6.28318530717958647692;
    DevRng rng = load_rng(rng_state, a);
    double *w_out = wi + 3 * a;
    if (fabs(g) < 1e-4) {
        const double z = 1.0 - 2.0 * rng.next_f64();
        const double rr = sqrt(fmax(1.0 - z * z, 0.0));
        const double phi = kTau * rng.next_f64();
        w_out[0] = rr * cos(phi);
        w_out[1] = rr * sin(phi);
        w_out[2] = z;
        pdf[a] = 1.0 / (4.0 * 3.14159265358979323846);
        store_rng(rng, rng_state, a);
        return;
    }
    const double sv = 2.0 * rng.next_f64() - 1.0;
    const double sqr = (1.0 - g * g) / (1.0 + g * sv);
    const double cos_theta = rclamp_d((1.0 + g * g - sqr * sqr) / (2.0 * g), -1.0, 1.0);
    const double sin_theta = sqrt(fmax(1.0 - cos_theta * cos_theta, 0.0));
    const double phi = kTau * rng.next_f64();
    const double *wo_a = wo + 3 * a;
    const double nwo[3] = {-wo_a[0], -wo_a[1], -wo_a[2]};
    double fwd[3], w[3];
    vnormalize3(nwo, fwd);
    vnormalize3(fwd, w);
    const double sign = copysign(1.0, w[2]);
    const double ca = -1.0 / (sign + w[2]);
    const double b = w[0] * w[1] * ca;
    const double u[3] = {1.0 + sign * w[0] * w[0] * ca, sign * b, -sign * w[0]};
    const double v[3] = {b, sign + w[1] * w[1] * ca, -w[1]};
    const double lx = sin_theta * cos(phi), ly = sin_theta * sin(phi), lz = cos_theta;
    const double world[3] = {u[0] * lx + v[0] * ly + w[0] * lz, u[1] * lx + v[1] * ly + w[1] * lz,
                             u[2] * lx + v[2] * ly + w[2] * lz};
    vnormalize3(world, w_out);
    pdf[a] = phase_hg(g, wo_a, w_out);
    store_rng(rng, rng_state, a);
}

}  // namespace
}  // namespace aether
