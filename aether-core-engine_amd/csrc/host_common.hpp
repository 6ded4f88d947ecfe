// host_common.hpp -- error handling, device buffers, launch-dispatch macros, parameter
// defaults and checks, and the kernel timer shared by the solver and the batch.
// Part of the library's single translation unit: included once, by
// fluid_api.hip, after the headers before it there.
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
#include "fluid_kernels.hpp"

namespace aether {
namespace {

thread_local std::string g_last_error;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string &msg) { throw Failure{code, msg}; }

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

// Instantiates a launch for the z planes per workgroup of the one-launch CG.
#define KZ_DISPATCH(kz, launch)                 \
    do {                                        \
        switch (kz) {                           \
        case 8: {                               \
            constexpr int KZ = 8;               \
            launch;                             \
        } break;                                \
        case 32: {                              \
            constexpr int KZ = 32;              \
            launch;                             \
        } break;                                \
        default: {                              \
            constexpr int KZ = 16;              \
            launch;                             \
        } break;                                \
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

aether_fluid_params default_params() {
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

void check_params(const aether_fluid_params &p) {
    if (p.solver_kind != AETHER_SOLVER_CG && p.solver_kind != AETHER_SOLVER_RBGS)
        raise(AETHER_EINVAL, "solver_kind must be AETHER_SOLVER_CG or AETHER_SOLVER_RBGS");
    if (p.solver_iters < 0) raise(AETHER_EINVAL, "solver_iters must be >= 0");
    if (p.cg_form < AETHER_CG_AUTO || p.cg_form > AETHER_CG_FUSED)
        raise(AETHER_EINVAL, "cg_form must be one of AETHER_CG_AUTO/SCALAR/X4/FUSED");
    if (p.source_kind < AETHER_SOURCE_JETS || p.source_kind > AETHER_SOURCE_NONE)
        raise(AETHER_EINVAL, "unknown source_kind");
    if (!(p.cg_rel_tol >= 0.0)) raise(AETHER_EINVAL, "cg_rel_tol must be >= 0");
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
    double sum_ms[4] = {0, 0, 0, 0};  // cg_dir, cg_update, cg_resident, cg_fused
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

}  // namespace
}  // namespace aether
