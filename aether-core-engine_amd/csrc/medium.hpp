// medium.hpp -- FluidMedium (src/fluid.rs:687-788) on the device.
// Included by the solver translation units (the medium samples the solver's
// storage type).
#pragma once
#include "host_common.hpp"
#include "fluid_kernels.hpp"
#include "volume_kernels.hpp"

namespace aether {
namespace {

// Device scratch of the medium's queries: one block, grown when a call needs
// more and reused by later calls (each call waits for its copies before it
// returns, and a handle is not shared between threads).  Sub-buffers start
// on 256-byte boundaries.
struct Arena {
    DevBuf<unsigned char> buf;
    size_t off = 0;
    static size_t round_up(size_t b) { return (b + 255) & ~size_t(255); }
    template <typename... B>
    void reserve(B... bytes) {
        const size_t need = (round_up(bytes) + ...);
        if (buf.n < need) buf.alloc(need);
        off = 0;
    }
    template <typename U>
    U *take(size_t count) {
        U *q = reinterpret_cast<U *>(buf.p + off);
        off += round_up(count * sizeof(U));
        return q;
    }
};

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
        scratch_.reserve(npts * 3 * sizeof(double), npts * 2 * sizeof(double));
        double *dxyz = scratch_.take<double>(npts * 3), *dout = scratch_.take<double>(npts * 2);
        HIPT(hipMemcpyAsync(dxyz, xyz, npts * 3 * sizeof(double), hipMemcpyHostToDevice, st_));
        k_medium_sample<T><<<(unsigned)((npts + kThreads - 1) / kThreads), kThreads, 0, st_>>>(
            g_, box_, rho_.p, temp_.p, dxyz, (long)npts, dout);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(out, dout, npts * 2 * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

    MediumDev dev() const {
        MediumDev m;
        m.g = g_;
        m.box = box_;
        m.sigma_t_scale = sigma_t_scale;
        m.max_density = max_density;
        m.majorant = majorant_sigma_t;
        m.phase_g = phase_g;
        return m;
    }
    static unsigned blocks(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

    // fluid.rs:793-838 for a batch of rays (inputs copied in, outputs and
    // the advanced RNG states copied back)
    void interaction(const aether_medium_ray *rays, size_t n, const double *lambda,
                     uint64_t *rng, aether_medium_interaction *out) override {
        if (n == 0) return;
        scratch_.reserve(n * sizeof(*rays), 4 * n * sizeof(double), 4 * n * sizeof(uint64_t),
                         n * sizeof(*out));
        auto *dr = scratch_.take<aether_medium_ray>(n);
        double *dl = scratch_.take<double>(4 * n);
        uint64_t *ds = scratch_.take<uint64_t>(4 * n);
        auto *dout = scratch_.take<aether_medium_interaction>(n);
        HIPT(hipMemcpyAsync(dr, rays, n * sizeof(*rays), hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(dl, lambda, 4 * n * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(ds, rng, 4 * n * sizeof(uint64_t), hipMemcpyHostToDevice, st_));
        k_medium_interaction<T><<<blocks(n), kThreads, 0, st_>>>(dev(), rho_.p, temp_.p, dr, dl,
                                                                 (long)n, ds, dout);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(out, dout, n * sizeof(*out), hipMemcpyDeviceToHost, st_));
        HIPT(hipMemcpyAsync(rng, ds, 4 * n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

    // fluid.rs:841-874
    void transmittance(const aether_medium_ray *rays, size_t n, uint64_t *rng,
                       double *out) override {
        if (n == 0) return;
        scratch_.reserve(n * sizeof(*rays), 4 * n * sizeof(uint64_t), n * sizeof(double));
        auto *dr = scratch_.take<aether_medium_ray>(n);
        uint64_t *ds = scratch_.take<uint64_t>(4 * n);
        double *dout = scratch_.take<double>(n);
        HIPT(hipMemcpyAsync(dr, rays, n * sizeof(*rays), hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(ds, rng, 4 * n * sizeof(uint64_t), hipMemcpyHostToDevice, st_));
        k_medium_transmittance<T><<<blocks(n), kThreads, 0, st_>>>(dev(), rho_.p, temp_.p, dr,
                                                                   (long)n, ds, dout);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipMemcpyAsync(rng, ds, 4 * n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

    // fluid.rs:878-888
    void phase_pdf(const double *wo, const double *wi, size_t n, double *pdf) override {
        if (n == 0) return;
        scratch_.reserve(3 * n * sizeof(double), 3 * n * sizeof(double), n * sizeof(double));
        double *dwo = scratch_.take<double>(3 * n), *dwi = scratch_.take<double>(3 * n);
        double *dp = scratch_.take<double>(n);
        HIPT(hipMemcpyAsync(dwo, wo, 3 * n * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(dwi, wi, 3 * n * sizeof(double), hipMemcpyHostToDevice, st_));
        k_phase_pdf<<<blocks(n), kThreads, 0, st_>>>(phase_g, dwo, dwi, (long)n, dp);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(pdf, dp, n * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

    // fluid.rs:892-910
    void phase_sample(const double *wo, size_t n, uint64_t *rng, double *wi,
                      double *pdf) override {
        if (n == 0) return;
        scratch_.reserve(3 * n * sizeof(double), 3 * n * sizeof(double), n * sizeof(double),
                         4 * n * sizeof(uint64_t));
        double *dwo = scratch_.take<double>(3 * n), *dwi = scratch_.take<double>(3 * n);
        double *dp = scratch_.take<double>(n);
        uint64_t *ds = scratch_.take<uint64_t>(4 * n);
        HIPT(hipMemcpyAsync(dwo, wo, 3 * n * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(ds, rng, 4 * n * sizeof(uint64_t), hipMemcpyHostToDevice, st_));
        k_phase_sample<<<blocks(n), kThreads, 0, st_>>>(phase_g, dwo, (long)n, ds, dwi, dp);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(wi, dwi, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipMemcpyAsync(pdf, dp, n * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipMemcpyAsync(rng, ds, 4 * n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

   private:
    Dims g_;
    Box3 box_;
    DevBuf<T> rho_, temp_;
    Arena scratch_;
    hipStream_t st_ = nullptr;
};

}  // namespace
}  // namespace aether
