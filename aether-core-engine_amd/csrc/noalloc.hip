// noalloc.hip -- telemetry of caller-held device fields with no device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The same two kernels as FluidSolver3D::compute_telemetry
// (src/fluid.rs:613-672; k_telemetry_part / k_telemetry_final), on the same
// reduction tiling, so the frame equals aether_fluid_compute_telemetry's
// bit for bit.  Their partial sums live in `scratch`, a file-scope device
// array every GPU holds once for the life of the process, instead of a
// buffer the solver allocates.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "aether_fluid.h"
#include "fluid_kernels.hpp"
#include "host_common.hpp"

namespace aether {
namespace {

// Partial-sum slots of the scratch array: one per workgroup of the
// reduction tiling, kTelParts doubles each.  A grid needs
// max(cols, ~1024) slots with cols = ceil(nx/64) * ceil(ny/4), so every
// grid with at most 4096 such columns fits (1024 x 1024 x any nz); larger
// ones are rejected.
constexpr int kNoAllocMaxSlots = 4096;
constexpr size_t kScratchDoubles = (size_t)kNoAllocMaxSlots * kTelParts;  // 224 KiB

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device.  Host code never passes the array's own name to a kernel: on the
// host it is only a stand-in for the symbol.
double *scratch_on_current_device() {
    int dev = 0;
    HIPT(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) raise(AETHER_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPT(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) raise(AETHER_EHIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

template <typename T>
void launch_telemetry(const RedTiling &rt, Dims g, double h, const void *u, const void *v,
                      const void *w, const void *rho, double t, double *part, double *out,
                      hipStream_t st) {
    k_telemetry_part<T><<<dim3(rt.gx, rt.gy, rt.nch), dim3(kBX, kBY, 1), 0, st>>>(
        g, rt.zchunk, h, static_cast<const T *>(u), static_cast<const T *>(v),
        static_cast<const T *>(w), static_cast<const T *>(rho), part, nullptr);
    LAUNCH_CHECK();
    k_telemetry_final<<<1, dim3(kBX, kBY, 1), 0, st>>>(part, rt.nb, t, out);
    LAUNCH_CHECK();
}

}  // namespace
}  // namespace aether

using namespace aether;

extern "C" int aether_fluid_compute_telemetry_noalloc(int precision, int nx, int ny, int nz,
                                                      double h, const void *u, const void *v,
                                                      const void *w, const void *density,
                                                      double time_sec, double *out_dev,
                                                      void *stream) {
    return guarded([&] {
        if (precision != AETHER_FP32 && precision != AETHER_FP64)
            raise(AETHER_EINVAL, "precision must be AETHER_FP32 or AETHER_FP64");
        if (nx < 6 || ny < 6 || nz < 6)
            raise(AETHER_EINVAL, "grid dimensions must be >= 6 (src/fluid.rs:54)");
        if ((double)nx * ny * nz >= 2147483647.0)
            raise(AETHER_EINVAL, "grid has more than 2^31-1 cells");
        if (!u || !v || !w || !density || !out_dev)
            raise(AETHER_EINVAL, "null field or output pointer");
        const RedTiling rt = red_tiling(nx, ny, nz);
        if (rt.nb > kNoAllocMaxSlots)
            raise(AETHER_EINVAL, "grid needs " + std::to_string(rt.nb) +
                                     " reduction partials; the no-allocation entry point holds " +
                                     std::to_string(kNoAllocMaxSlots));
        // src/fluid.rs:56 when h <= 0
        const double hh = h > 0.0 ? h : 1.0 / (double)std::max(nx, std::max(ny, nz));
        double *part = scratch_on_current_device();
        Dims g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.sxy = nx * ny;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == AETHER_FP64)
            launch_telemetry<double>(rt, g, hh, u, v, w, density, time_sec, part, out_dev, st);
        else
            launch_telemetry<float>(rt, g, hh, u, v, w, density, time_sec, part, out_dev, st);
    });
}
