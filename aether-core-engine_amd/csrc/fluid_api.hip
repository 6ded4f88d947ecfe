// fluid_api.hip -- host side of the MI355X smoke solver behind the C ABI of
// include/aether_fluid.h.
//
// One handle = one grid, one HIP stream and its device buffers.  All state
// stays resident in HBM; fields cross to the host only through
// upload/download.  The stage sequence of step() follows
// FluidSolver3D::step (src/fluid.rs:598-611); the CG driver keeps every
// scalar (alpha, beta, r.r, the stop flag) on the device, so a solve is a
// fixed sequence of launches with no host round trip.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "aether_fluid.h"
#include "host_common.hpp"
#include "fluid_kernels.hpp"


using namespace aether;

struct aether_fluid_batch_s {
    int device;
    hipStream_t stream = nullptr;
    std::unique_ptr<BatchBase> impl;
    ~aether_fluid_batch_s() {
        impl.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct aether_fluid_s {
    int nx, ny, nz, precision;
    double h;
    int device;
    hipStream_t stream = nullptr;
    std::unique_ptr<SolverBase> impl;
    ~aether_fluid_s() {
        impl.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
aether_fluid_s *checked(aether_fluid_t s) {
    if (!s) raise(AETHER_EINVAL, "null solver handle");
    HIPT(hipSetDevice(s->device));
    return s;
}
size_t cells(const aether_fluid_s *s) { return (size_t)s->nx * s->ny * s->nz; }

// SplitMix64 seeding and Xoshiro256++ (src/rng.rs:5-64).
uint64_t splitmix64(uint64_t &st) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
uint64_t xoshiro_next(uint64_t s[4]) {
    const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}
}  // namespace

extern "C" {

int aether_fluid_abi_version(void) { return AETHER_FLUID_ABI_VERSION; }

const char *aether_fluid_last_error(void) { return g_last_error.c_str(); }

int aether_fluid_device_count(int *count) {
    return guarded([&] {
        if (!count) raise(AETHER_EINVAL, "null count pointer");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = e == hipSuccess ? c : 0;
    });
}

int aether_fluid_create(int nx, int ny, int nz, double h, int precision, aether_fluid_t *out) {
    return guarded([&] {
        if (!out) raise(AETHER_EINVAL, "null output handle pointer");
        *out = nullptr;
        if (nx < 6 || ny < 6 || nz < 6)
            raise(AETHER_EINVAL, "grid dimensions must be >= 6 (src/fluid.rs:54)");
        if ((double)nx * ny * nz >= 2147483647.0)
            raise(AETHER_EINVAL, "grid has more than 2^31-1 cells");
        if (precision != AETHER_FP32 && precision != AETHER_FP64 && precision != AETHER_FP32_IR64)
            raise(AETHER_EINVAL, "precision must be AETHER_FP32, AETHER_FP64 or AETHER_FP32_IR64");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            raise(AETHER_ENODEV, "no HIP device available");
        auto s = std::make_unique<aether_fluid_s>();
        s->nx = nx;
        s->ny = ny;
        s->nz = nz;
        s->precision = precision;
        // src/fluid.rs:56
        s->h = h > 0.0 ? h : 1.0 / (double)std::max(nx, std::max(ny, nz));
        HIPT(hipGetDevice(&s->device));
        HIPT(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        if (precision == AETHER_FP64)
            s->impl = make_solver_f64(nx, ny, nz, s->h, s->stream);
        else
            s->impl = make_solver_f32(nx, ny, nz, s->h, s->stream, precision == AETHER_FP32_IR64);
        *out = s.release();
    });
}

int aether_fluid_destroy(aether_fluid_t s) {
    return guarded([&] {
        if (!s) return;
        (void)hipSetDevice(s->device);
        delete s;
    });
}

int aether_fluid_dims(aether_fluid_t s, int *nx, int *ny, int *nz, double *h, int *precision) {
    return guarded([&] {
        if (!s) raise(AETHER_EINVAL, "null solver handle");
        if (nx) *nx = s->nx;
        if (ny) *ny = s->ny;
        if (nz) *nz = s->nz;
        if (h) *h = s->h;
        if (precision) *precision = s->precision;
    });
}

int aether_fluid_default_params(aether_fluid_params *out) {
    return guarded([&] {
        if (!out) raise(AETHER_EINVAL, "null params pointer");
        *out = default_params();
    });
}

int aether_fluid_set_params(aether_fluid_t s, const aether_fluid_params *p) {
    return guarded([&] {
        checked(s);
        if (!p) raise(AETHER_EINVAL, "null params pointer");
        check_params(*p);
        s->impl->params = *p;
    });
}

int aether_fluid_get_params(aether_fluid_t s, aether_fluid_params *p) {
    return guarded([&] {
        checked(s);
        if (!p) raise(AETHER_EINVAL, "null params pointer");
        *p = s->impl->params;
    });
}

int aether_fluid_step(aether_fluid_t s, double dt, double time_sec, aether_fluid_telemetry *out) {
    return guarded([&] {
        checked(s);
        aether_fluid_telemetry tmp;
        s->impl->run(dt, &time_sec, 1, &tmp);
        if (out) *out = tmp;
    });
}

int aether_fluid_run(aether_fluid_t s, double dt, const double *times, int nsteps,
                     aether_fluid_telemetry *out) {
    return guarded([&] {
        checked(s);
        if (nsteps > 0 && !times) raise(AETHER_EINVAL, "null times array");
        s->impl->run(dt, times, nsteps, out);
    });
}

int aether_fluid_advect_maccormack_in_place(aether_fluid_t s, int field_kind, double dt, double dissipation) {
    return guarded([&] { checked(s)->impl->advect(field_kind, dt, dissipation); });
}

int aether_fluid_enforce_wall_velocity_bc(aether_fluid_t s) {
    return guarded([&] { checked(s)->impl->enforce_bc(); });
}

int aether_fluid_apply_buoyancy_and_vorticity_confinement(aether_fluid_t s, double dt) {
    return guarded([&] { checked(s)->impl->buoyancy_confinement(dt); });
}

int aether_fluid_compute_divergence_norm(aether_fluid_t s, double *rms) {
    return guarded([&] {
        checked(s);
        if (!rms) raise(AETHER_EINVAL, "null output pointer");
        *rms = s->impl->divergence_norm();
    });
}

int aether_fluid_divergence_max(aether_fluid_t s, double *out) {
    return guarded([&] {
        checked(s);
        if (!out) raise(AETHER_EINVAL, "null output pointer");
        *out = s->impl->divergence_max();
    });
}

int aether_fluid_project(aether_fluid_t s) {
    return guarded([&] { checked(s)->impl->project(); });
}

int aether_fluid_solve_poisson_conjugate_gradient(aether_fluid_t s, int max_iters) {
    return guarded([&] { checked(s)->impl->solve_cg(max_iters); });
}

int aether_fluid_solve_poisson_gauss_seidel(aether_fluid_t s, int sweeps) {
    return guarded([&] { checked(s)->impl->solve_rbgs(sweeps); });
}

int aether_fluid_inject_sources(aether_fluid_t s, double time_sec, double dt) {
    return guarded([&] { checked(s)->impl->inject_sources(time_sec, dt); });
}

int aether_fluid_compute_telemetry(aether_fluid_t s, double time_sec, aether_fluid_telemetry *out) {
    return guarded([&] {
        checked(s);
        if (!out) raise(AETHER_EINVAL, "null output pointer");
        s->impl->telemetry(time_sec, out);
    });
}

int aether_fluid_sample_trilinear(aether_fluid_t s, int field, const double *xyz, size_t npts,
                                  double *out) {
    return guarded([&] {
        checked(s);
        if (npts && (!xyz || !out)) raise(AETHER_EINVAL, "null point or output array");
        s->impl->sample(field, xyz, npts, out);
    });
}

int aether_fluid_sample_trilinear_field(aether_fluid_t s, const double *field, size_t n,
                                        const double *xyz, size_t npts, double *out) {
    return guarded([&] {
        checked(s);
        if (!field) raise(AETHER_EINVAL, "null field array");
        if (n != cells(s))
            raise(AETHER_EINVAL, "field size " + std::to_string(n) + " != nx*ny*nz = " +
                                     std::to_string(cells(s)));
        if (npts && (!xyz || !out)) raise(AETHER_EINVAL, "null point or output array");
        s->impl->sample_field(field, xyz, npts, out);
    });
}

int aether_fluid_last_solve_stats(aether_fluid_t s, aether_fluid_solve_stats *out) {
    return guarded([&] {
        checked(s);
        if (!out) raise(AETHER_EINVAL, "null output pointer");
        s->impl->stats(out);
    });
}

int aether_fluid_last_run_iterations(aether_fluid_t s, int32_t *out, int n, int *count) {
    return guarded([&] {
        checked(s);
        if (n > 0 && !out) raise(AETHER_EINVAL, "null output array");
        const auto &v = s->impl->last_run_iters();
        if (count) *count = (int)v.size();
        for (int a = 0; a < n && a < (int)v.size(); ++a) out[a] = v[a];
    });
}

int aether_fluid_nonfinite(aether_fluid_t s, int *flag) {
    return guarded([&] {
        checked(s);
        if (!flag) raise(AETHER_EINVAL, "null flag pointer");
        *flag = s->impl->take_nonfinite();
    });
}

int aether_fluid_upload(aether_fluid_t s, int field, const double *src, size_t n) {
    return guarded([&] {
        checked(s);
        if (n != cells(s))
            raise(AETHER_EINVAL, "upload size " + std::to_string(n) + " != nx*ny*nz = " +
                                     std::to_string(cells(s)));
        if (!src) raise(AETHER_EINVAL, "null source array");
        s->impl->upload(field, src);
    });
}

int aether_fluid_download(aether_fluid_t s, int field, double *dst, size_t n) {
    return guarded([&] {
        checked(s);
        if (n != cells(s))
            raise(AETHER_EINVAL, "download size " + std::to_string(n) + " != nx*ny*nz = " +
                                     std::to_string(cells(s)));
        if (!dst) raise(AETHER_EINVAL, "null destination array");
        s->impl->download(field, dst);
    });
}

int aether_fluid_max_density(aether_fluid_t s, double *out) {
    return guarded([&] {
        checked(s);
        if (!out) raise(AETHER_EINVAL, "null output pointer");
        *out = s->impl->max_density();
    });
}

int aether_fluid_batch_autotune_cg(aether_fluid_batch_t b, int iters, double *us_per_iter,
                                   int *form, int *z_cells) {
    return guarded([&] {
        if (!b) raise(AETHER_EINVAL, "null batch handle");
        HIPT(hipSetDevice(b->device));
        b->impl->autotune(iters, us_per_iter, form, z_cells);
    });
}

int aether_fluid_autotune_cg(aether_fluid_t s, int iters, double *us_per_iter, int *form,
                             int *z_cells) {
    return guarded([&] {
        checked(s);
        s->impl->autotune(iters, us_per_iter, form, z_cells);
    });
}

struct aether_fluid_medium_s {
    int device;
    std::unique_ptr<MediumBase> impl;
};

int aether_fluid_medium_create(aether_fluid_t s, const double *bounds_min,
                               const double *bounds_max, double sigma_t_scale, double phase_g,
                               aether_fluid_medium_t *out) {
    return guarded([&] {
        if (!out) raise(AETHER_EINVAL, "null output handle pointer");
        *out = nullptr;
        checked(s);
        if (!bounds_min || !bounds_max) raise(AETHER_EINVAL, "null bounds");
        auto m = std::make_unique<aether_fluid_medium_s>();
        m->device = s->device;
        m->impl = s->impl->make_medium(bounds_min, bounds_max);
        // src/fluid.rs:709, 720
        m->impl->sigma_t_scale = sigma_t_scale;
        m->impl->majorant_sigma_t = std::max(m->impl->max_density * sigma_t_scale, 1e-4);
        m->impl->phase_g = std::min(std::max(phase_g, -0.9), 0.9);
        *out = m.release();
    });
}

int aether_fluid_medium_destroy(aether_fluid_medium_t m) {
    return guarded([&] {
        if (!m) return;
        (void)hipSetDevice(m->device);
        delete m;
    });
}

int aether_fluid_medium_info(aether_fluid_medium_t m, double *max_density,
                             double *majorant_sigma_t, double *phase_g) {
    return guarded([&] {
        if (!m) raise(AETHER_EINVAL, "null medium handle");
        if (max_density) *max_density = m->impl->max_density;
        if (majorant_sigma_t) *majorant_sigma_t = m->impl->majorant_sigma_t;
        if (phase_g) *phase_g = m->impl->phase_g;
    });
}

int aether_fluid_medium_sample(aether_fluid_medium_t m, const double *xyz, size_t npts,
                               double *out) {
    return guarded([&] {
        if (!m) raise(AETHER_EINVAL, "null medium handle");
        if (npts && (!xyz || !out)) raise(AETHER_EINVAL, "null point or output array");
        HIPT(hipSetDevice(m->device));
        m->impl->sample(xyz, npts, out);
    });
}

namespace {
aether_fluid_medium_s *checked_medium(aether_fluid_medium_t m) {
    if (!m) raise(AETHER_EINVAL, "null medium handle");
    HIPT(hipSetDevice(m->device));
    return m;
}
}  // namespace

int aether_fluid_medium_sample_interaction(aether_fluid_medium_t m, const aether_medium_ray *rays,
                                           size_t nrays, const double *lambda,
                                           uint64_t *rng_state, aether_medium_interaction *out) {
    return guarded([&] {
        checked_medium(m);
        if (nrays && (!rays || !lambda || !rng_state || !out))
            raise(AETHER_EINVAL, "null ray, wavelength, rng or output array");
        m->impl->interaction(rays, nrays, lambda, rng_state, out);
    });
}

int aether_fluid_medium_transmittance(aether_fluid_medium_t m, const aether_medium_ray *rays,
                                      size_t nrays, uint64_t *rng_state, double *out) {
    return guarded([&] {
        checked_medium(m);
        if (nrays && (!rays || !rng_state || !out))
            raise(AETHER_EINVAL, "null ray, rng or output array");
        m->impl->transmittance(rays, nrays, rng_state, out);
    });
}

int aether_fluid_medium_phase_hg_pdf(aether_fluid_medium_t m, const double *wo, const double *wi,
                                     size_t n, double *pdf) {
    return guarded([&] {
        checked_medium(m);
        if (n && (!wo || !wi || !pdf)) raise(AETHER_EINVAL, "null direction or output array");
        m->impl->phase_pdf(wo, wi, n, pdf);
    });
}

int aether_fluid_medium_sample_phase_hg(aether_fluid_medium_t m, const double *wo, size_t n,
                                        uint64_t *rng_state, double *wi, double *pdf) {
    return guarded([&] {
        checked_medium(m);
        if (n && (!wo || !rng_state || !wi || !pdf))
            raise(AETHER_EINVAL, "null direction, rng or output array");
        m->impl->phase_sample(wo, n, rng_state, wi, pdf);
    });
}

int aether_fluid_batch_create(const aether_fluid_instance_desc *desc, int count, int precision,
                              const aether_fluid_params *shared, aether_fluid_batch_t *out) {
    return guarded([&] {
        if (!out || !desc || !shared) raise(AETHER_EINVAL, "null argument");
        *out = nullptr;
        if (count < 1) raise(AETHER_EINVAL, "count must be >= 1");
        if (precision != AETHER_FP32 && precision != AETHER_FP64)
            raise(AETHER_EINVAL, "batched precision must be AETHER_FP32 or AETHER_FP64");
        check_params(*shared);
        if (shared->solver_kind != AETHER_SOLVER_CG)
            raise(AETHER_EINVAL, "batched instances use the CG projection");
        for (int q = 0; q < count; ++q) {
            const auto &d = desc[q];
            if (d.nx < 6 || d.ny < 6 || d.nz < 6)
                raise(AETHER_EINVAL, "instance " + std::to_string(q) + ": dimensions must be >= 6");
            if ((double)d.nx * d.ny * d.nz >= 2147483647.0)
                raise(AETHER_EINVAL, "instance grid has more than 2^31-1 cells");
            if (!(std::isfinite(d.source_center[0]) && std::isfinite(d.source_center[1]) &&
                  std::isfinite(d.source_center[2]) && std::isfinite(d.source_radius) &&
                  std::fabs(d.source_radius) < 1e6))
                raise(AETHER_EINVAL, "instance source must be finite");
        }
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            raise(AETHER_ENODEV, "no HIP device available");
        auto b = std::make_unique<aether_fluid_batch_s>();
        HIPT(hipGetDevice(&b->device));
        HIPT(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->impl = make_batch(desc, count, precision, *shared, b->stream);
        *out = b.release();
    });
}

int aether_fluid_batch_destroy(aether_fluid_batch_t b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->device);
        delete b;
    });
}

namespace {
aether_fluid_batch_s *checked_batch(aether_fluid_batch_t b) {
    if (!b) raise(AETHER_EINVAL, "null batch handle");
    HIPT(hipSetDevice(b->device));
    return b;
}
}  // namespace

int aether_fluid_batch_upload(aether_fluid_batch_t b, int inst, int field, const double *src,
                              size_t n) {
    return guarded([&] {
        if (!src) raise(AETHER_EINVAL, "null source array");
        checked_batch(b)->impl->upload(inst, field, src, n);
    });
}

int aether_fluid_batch_download(aether_fluid_batch_t b, int inst, int field, double *dst,
                                size_t n) {
    return guarded([&] {
        if (!dst) raise(AETHER_EINVAL, "null destination array");
        checked_batch(b)->impl->download(inst, field, dst, n);
    });
}

int aether_fluid_batch_run(aether_fluid_batch_t b, double dt, const double *times, int nsteps,
                           aether_fluid_telemetry *out) {
    return guarded([&] {
        checked_batch(b);
        if (nsteps < 0) raise(AETHER_EINVAL, "nsteps must be >= 0");
        if (nsteps > 0 && !times) raise(AETHER_EINVAL, "null times array");
        b->impl->run(dt, times, nsteps, out);
    });
}

int aether_fluid_batch_iterations(aether_fluid_batch_t b, int32_t *out, int n, int *count) {
    return guarded([&] {
        checked_batch(b);
        if (n > 0 && !out) raise(AETHER_EINVAL, "null output array");
        const auto &v = b->impl->iters();
        if (count) *count = (int)v.size();
        for (int a = 0; a < n && a < (int)v.size(); ++a) out[a] = v[a];
    });
}

int aether_fluid_rollout(aether_fluid_t s, const aether_rollout_desc *d, const double *controls,
                         double beta, const double *target, double *loss, double *grad_controls,
                         double *grad_beta) {
    return guarded([&] {
        checked(s);
        if (!d || !controls || !target || !loss) raise(AETHER_EINVAL, "null argument");
        *loss = s->impl->rollout(*d, controls, beta, target, grad_controls, grad_beta);
    });
}

int aether_fluid_rollout_forward(aether_fluid_t s, const aether_rollout_desc *d,
                                 const double *controls, double beta, double *final_state) {
    return guarded([&] {
        checked(s);
        if (!d || !controls) raise(AETHER_EINVAL, "null argument");
        s->impl->rollout_forward(*d, controls, beta, final_state);
    });
}

int aether_fluid_rollout_backward(aether_fluid_t s, const double *lam_final,
                                  double *grad_controls, double *grad_beta,
                                  double *grad_initial) {
    return guarded([&] {
        checked(s);
        s->impl->rollout_backward(lam_final, grad_controls, grad_beta, grad_initial);
    });
}

int aether_rng_seed(uint64_t seed, uint64_t state[4]) {
    return guarded([&] {
        if (!state) raise(AETHER_EINVAL, "null state");
        uint64_t sm = seed ^ 0xA5A55A5AF0F00F0Full;  // src/rng.rs:22
        for (int a = 0; a < 4; ++a) state[a] = splitmix64(sm);
        if (!state[0] && !state[1] && !state[2] && !state[3]) state[0] = 1;
    });
}

int aether_rng_fill_range(uint64_t state[4], double lo, double hi, double *out, size_t n) {
    return guarded([&] {
        if (!state || (n && !out)) raise(AETHER_EINVAL, "null state or output");
        for (size_t a = 0; a < n; ++a) {
            // src/rng.rs:74-77, 86-88
            const double f = (double)(xoshiro_next(state) >> 11) * (1.0 / 9007199254740992.0);
            out[a] = lo + (hi - lo) * f;
        }
    });
}

int aether_fluid_stream_copy(size_t nbytes, int reps, double *avg_ms) {
    return guarded([&] {
        if (nbytes < 16 || nbytes % 16 || reps < 1 || !avg_ms)
            raise(AETHER_EINVAL, "nbytes must be a positive multiple of 16, reps >= 1");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            raise(AETHER_ENODEV, "no HIP device available");
        hipStream_t st;
        HIPT(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct Guard {
            hipStream_t s;
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Guard() {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
                (void)hipStreamDestroy(s);
            }
        } guard{st};
        DevBuf<float4> a(nbytes / 16), b(nbytes / 16);
        const long n4 = (long)(nbytes / 16);
        int dev = 0, cus = 256;
        HIPT(hipGetDevice(&dev));
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const unsigned blocks = (unsigned)std::min<long>((n4 + kThreads - 1) / kThreads, 8L * cus);
        k_fill_hash<float><<<blocks, kThreads, 0, st>>>(reinterpret_cast<float *>(a.p), 4 * n4);
        LAUNCH_CHECK();
        k_fill_hash<float><<<blocks, kThreads, 0, st>>>(reinterpret_cast<float *>(b.p), 4 * n4);
        LAUNCH_CHECK();
        k_stream_copy<<<blocks, kThreads, 0, st>>>(a.p, n4, b.p);  // warm-up
        LAUNCH_CHECK();
        HIPT(hipEventCreate(&guard.e[0]));
        HIPT(hipEventCreate(&guard.e[1]));
        hipEvent_t e0 = guard.e[0], e1 = guard.e[1];
        HIPT(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) {
            k_stream_copy<<<blocks, kThreads, 0, st>>>(a.p, n4, b.p);
            LAUNCH_CHECK();
        }
        HIPT(hipEventRecord(e1, st));
        HIPT(hipEventSynchronize(e1));
        float ms = 0;
        HIPT(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = ms / reps;
    });
}

int aether_fluid_set_kernel_timing(aether_fluid_t s, int enable) {
    return guarded([&] {
        checked(s);
        s->impl->timer.enabled = enable != 0;
        s->impl->timer.reset();
    });
}

int aether_fluid_kernel_timing(aether_fluid_t s, int which, double *avg_ms, long *launches) {
    return guarded([&] {
        checked(s);
        if (which < 0 || which > 3 || which == 2)
            raise(AETHER_EINVAL, "which must be 0 (cg_dir), 1 (cg_update) or 3 (cg_fused)");
        const auto &t = s->impl->timer;
        if (avg_ms) *avg_ms = t.count[which] ? t.sum_ms[which] / t.count[which] : 0.0;
        if (launches) *launches = t.count[which];
    });
}

}  // extern "C"
