// host_common.hpp -- error handling, device buffers, launch-dispatch macros, parameter
// defaults and checks, and the kernel timer shared by the solver and the batch.
// Shared by every translation unit of the library: everything here has
// external linkage (inline), so an error raised in one unit is caught in
// another.
#pragma once
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

namespace aether {

inline thread_local std::string g_last_error;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] inline void raise(int code, const std::string &msg) { throw Failure{code, msg}; }

#define HIPT(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            raise(e_ == hipErrorOutOfMemory ? AETHER_ENOMEM : AETHER_EHIP,                \
                  std::string(#expr) + ": " + hipGetErrorString(e_) + " (line " +         \
                      std::to_string(__LINE__) + ")");                                    \
    } while (0)

#define LAUNCH_CHECK() HIPT(hipGetLastError())

// Instantiates a launch for the z cells per thread of the four-wide CG tiling.
#define VZU_DISPATCH(zu, launch)                \
    do {                                        \
        if ((zu) == 2) {                        \
            constexpr int VZU = 2;              \
            launch;                             \
        } else {                                \
            constexpr int VZU = 4;              \
            launch;                             \
        }                                       \
    } while (0)

// Instantiates a launch for the z-cells-per-thread of the CG tiling.
#define ZU_DISPATCH(zu, launch)                 \
    do {                                        \
        switch (zu) {                           \
        case 4: {                               \
            constexpr int ZU = 4;               \
            launch;                             \
        } break;                                \
        case 8: {                               \
            constexpr int ZU = 8;               \
            launch;                             \
        } break;                                \
        default: {                              \
            constexpr int ZU = 16;              \
            launch;                             \
        } break;                                \
        }                                       \
    } while (0)

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return AETHER_OK;
    } catch (const Failure &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return AETHER_ENOMEM;
    } catch (const std::exception &e) {  // nothing may unwind through the C ABI
        g_last_error = std::string("internal error: ") + e.what();
        return AETHER_EINTERNAL;
    }
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        if (count) HIPT(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

inline aether_fluid_params default_params() {
    aether_fluid_params p;
    std::memset(&p, 0, sizeof(p));
    p.vorticity_eps = 5.5;  // src/fluid.rs:79-83
    p.buoyancy_beta = 6.8;
    p.gravity_alpha = 0.45;
    p.solver_kind = AETHER_SOLVER_CG;
    p.solver_iters = 32;
    p.sor_omega = 1.65;  // src/fluid.rs:398
    p.cg_rel_tol = 0.0;
    p.dissipation_velocity = 0.996;  // src/fluid.rs:599-607
    p.dissipation_density = 0.996;
    p.dissipation_temperature = 0.985;  // src/fluid.rs:608
    p.source_kind = AETHER_SOURCE_JETS;
    p.source_radius = 0.0;
    p.source_density = 1.0;
    p.source_temperature = 1.0;
    p.ir_inner_rel_tol = 1e-3;
    p.ir_div_tol = 1e-7;
    p.ir_max_outer = 30;
    return p;
}

inline void check_params(const aether_fluid_params &p) {
    if (p.solver_kind != AETHER_SOLVER_CG && p.solver_kind != AETHER_SOLVER_RBGS)
        raise(AETHER_EINVAL, "solver_kind must be AETHER_SOLVER_CG or AETHER_SOLVER_RBGS");
    if (p.solver_iters < 0) raise(AETHER_EINVAL, "solver_iters must be >= 0");
    if (p.cg_form < AETHER_CG_AUTO || p.cg_form > AETHER_CG_FUSED)
        raise(AETHER_EINVAL, "cg_form must be one of AETHER_CG_AUTO/SCALAR/X4/FUSED");
    if (p.source_kind < AETHER_SOURCE_JETS || p.source_kind > AETHER_SOURCE_NONE)
        raise(AETHER_EINVAL, "unknown source_kind");
    if (!(p.cg_rel_tol >= 0.0)) raise(AETHER_EINVAL, "cg_rel_tol must be >= 0");
    if (p.cg_z_cells != 0 && p.cg_z_cells != 2 && p.cg_z_cells != 4 && p.cg_z_cells != 8 &&
        p.cg_z_cells != 16)
        raise(AETHER_EINVAL, "cg_z_cells must be 0, 2, 4, 8 or 16");
    if (p.source_kind == AETHER_SOURCE_SPHERE &&
        !(std::isfinite(p.source_center[0]) && std::isfinite(p.source_center[1]) &&
          std::isfinite(p.source_center[2]) && std::isfinite(p.source_radius) &&
          std::fabs(p.source_center[0]) < 1e9 && std::fabs(p.source_center[1]) < 1e9 &&
          std::fabs(p.source_center[2]) < 1e9 && p.source_radius < 1e9))
        raise(AETHER_EINVAL, "sphere source centre/radius must be finite");
}

// Kernel timing accumulator (HIP events on the solver's stream).
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    struct Rec {
        int which, step, it;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    size_t used = 0;
    double sum_ms[4] = {0, 0, 0, 0};  // cg_dir, cg_update, (unused), cg_fused
    long count[4] = {0, 0, 0, 0};
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIPT(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void reset() {
        recs.clear();
        used = 0;
        for (int a = 0; a < 4; ++a) {
            sum_ms[a] = 0;
            count[a] = 0;
        }
    }
    ~KernelTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Interfaces shared across translation units; the implementations are
// class templates private to solver_f32.hip, solver_f64.hip and batch.hip.

// FluidMedium (src/fluid.rs:687-788): device snapshots of density and
// temperature taken at from_solver, sampled by sample_density_and_temp.
class MediumBase {
   public:
    virtual ~MediumBase() = default;
    virtual void sample(const double *xyz, size_t npts, double *out) = 0;
    virtual void interaction(const aether_medium_ray *rays, size_t n, const double *lambda,
                             uint64_t *rng, aether_medium_interaction *out) = 0;
    virtual void transmittance(const aether_medium_ray *rays, size_t n, uint64_t *rng,
                               double *out) = 0;
    virtual void phase_pdf(const double *wo, const double *wi, size_t n, double *pdf) = 0;
    virtual void phase_sample(const double *wo, size_t n, uint64_t *rng, double *wi,
                              double *pdf) = 0;
    double max_density = 0.0, majorant_sigma_t = 0.0, phase_g = 0.0, sigma_t_scale = 0.0;
};

// One grid (FluidSolver3D, src/fluid.rs:23-673).
class SolverBase {
   public:
    virtual ~SolverBase() = default;
    virtual void advect(int kind, double dt, double diss) = 0;
    virtual void enforce_bc() = 0;
    virtual void buoyancy_confinement(double dt) = 0;
    virtual double divergence_norm() = 0;
    virtual double divergence_max() = 0;
    virtual void project() = 0;
    virtual void solve_cg(int iters) = 0;
    virtual void solve_rbgs(int sweeps) = 0;
    virtual void inject_sources(double t, double dt) = 0;
    virtual void telemetry(double t, aether_fluid_telemetry *out) = 0;
    virtual void run(double dt, const double *times, int nsteps, aether_fluid_telemetry *out) = 0;
    virtual void upload(int field, const double *src) = 0;
    virtual void download(int field, double *dst) = 0;
    virtual void sample(int field, const double *xyz, size_t npts, double *out) = 0;
    virtual void sample_field(const double *field, const double *xyz, size_t npts,
                              double *out) = 0;
    virtual double max_density() = 0;
    virtual std::unique_ptr<MediumBase> make_medium(const double *lo, const double *hi) = 0;
    virtual void autotune(int iters, double *us, int *form, int *z) = 0;
    virtual void stats(aether_fluid_solve_stats *out) = 0;
    virtual const std::vector<int> &last_run_iters() const = 0;
    // 1 if a telemetry reduction since the last call met a non-finite u, v, w,
    // density or |curl|^2; clears the flag
    virtual int take_nonfinite() = 0;
    virtual double rollout(const aether_rollout_desc &d, const double *ctrl, double beta,
                           const double *target, double *grad_c, double *grad_beta) = 0;
    virtual void rollout_forward(const aether_rollout_desc &d, const double *ctrl, double beta,
                                 double *final_state) = 0;
    virtual void rollout_backward(const double *lam_final, double *grad_c, double *grad_beta,
                                  double *grad_init) = 0;
    aether_fluid_params params = default_params();
    KernelTimer timer;
};

// Many grids advanced together, one launch per stage.
class BatchBase {
   public:
    virtual ~BatchBase() = default;
    virtual void autotune(int iters, double *us, int *form, int *z) = 0;
    virtual void upload(int inst, int field, const double *src, size_t n) = 0;
    virtual void download(int inst, int field, double *dst, size_t n) = 0;
    virtual void run(double dt, const double *times, int nsteps, aether_fluid_telemetry *out) = 0;
    virtual const std::vector<int> &iters() const = 0;
    virtual int count() const = 0;
};

std::unique_ptr<SolverBase> make_solver_f32(int nx, int ny, int nz, double h, hipStream_t st,
                                            bool ir64);
std::unique_ptr<SolverBase> make_solver_f64(int nx, int ny, int nz, double h, hipStream_t st);
std::unique_ptr<BatchBase> make_batch(const aether_fluid_instance_desc *desc, int count,
                                      int precision, const aether_fluid_params &shared,
                                      hipStream_t st);

}  // namespace aether
