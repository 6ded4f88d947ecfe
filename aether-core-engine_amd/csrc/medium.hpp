// medium.hpp -- FluidMedium (src/fluid.rs:687-788) on the device.
// Included by the solver translation units (the medium samples the solver's
// storage type).
#pragma once
#include "host_common.hpp"
#include "fluid_kernels.hpp"

namespace aether {
namespace {

template <typename T>
class Medium final : public MediumBase {
   public:
    // Copies density and temperature (on `copy_stream`, which must be idle
    // afterwards) into buffers the medium owns; samples on its own stream.
    Medium(Dims g, const T *rho, const T *temp, const double *lo, const double *hi,
           hipStream_t copy_stream)
        : g_(g) {
        const size_t n = (size_t)g.sxy * g.nz;
        rho_.alloc(n);
        temp_.alloc(n);
        HIPT(hipMemcpyAsync(rho_.p, rho, n * sizeof(T), hipMemcpyDeviceToDevice, copy_stream));
        HIPT(hipMemcpyAsync(temp_.p, temp, n * sizeof(T), hipMemcpyDeviceToDevice, copy_stream));
        HIPT(hipStreamSynchronize(copy_stream));
        for (int a = 0; a < 3; ++a) {
            box_.lo[a] = lo[a];
            box_.hi[a] = hi[a];
        }
        HIPT(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    }
    ~Medium() override {
        if (st_) (void)hipStreamDestroy(st_);
    }
    void sample(const double *xyz, size_t npts, double *out) override {
        if (npts == 0) return;
        DevBuf<double> dxyz(npts * 3), dout(npts * 2);
        HIPT(hipMemcpyAsync(dxyz.p, xyz, npts * 3 * sizeof(double), hipMemcpyHostToDevice, st_));
        k_medium_sample<T><<<(unsigned)((npts + kThreads - 1) / kThreads), kThreads, 0, st_>>>(
            g_, box_, rho_.p, temp_.p, dxyz.p, (long)npts, dout.p);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(out, dout.p, npts * 2 * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

   private:
    Dims g_;
    Box3 box_;
    DevBuf<T> rho_, temp_;
    hipStream_t st_ = nullptr;
};

}  // namespace
}  // namespace aether
