// solver.hpp -- one grid: the smoke step, its stages, the pressure solvers, the
// autotuner and the differentiable rollout.
// Included by solver_f32.hip and solver_f64.hip, which instantiate it.
#pragma once
#include "host_common.hpp"
#include "medium.hpp"
#include "cg_fused.hpp"
#include "cg_vec.hpp"
#include "cg_vec2.hpp"
#include "cg_cgs4.hpp"
#include "adjoint_kernels.hpp"

namespace aether {
namespace {

template <typename T>
class Solver final : public SolverBase {
    // x-cells per thread of the 16-byte-lane CG layout (AETHER_CG_X4)
    static constexpr int kLaneCells = sizeof(T) == 4 ? 4 : 2;

   public:
    Solver(int nx, int ny, int nz, double h, hipStream_t stream, bool ir64 = false)
        : h_(h), st_(stream), ir64_(ir64) {
        g_.nx = nx;
        g_.ny = ny;
        g_.nz = nz;
        g_.sxy = nx * ny;
        n_ = (size_t)nx * ny * nz;
        cell_blk_ = dim3(kBX, kBY, 1);
        const int gx = (nx + kBX - 1) / kBX, gy = (ny + kBY - 1) / kBY;
        cell_grid_ = dim3(gx, gy, nz);
        // reduction launches: each workgroup marches a chunk of z planes so
        // that a solve sees ~1k partial sums at most
        const int cols = gx * gy;
        const RedTiling rt = red_tiling(nx, ny, nz);
        zchunk_ = rt.zchunk;
        red_grid_ = dim3(rt.gx, rt.gy, rt.nch);
        nb_ = rt.nb;
        // CG tiling: ZU z cells per thread; the smallest ZU giving <= 1024
        // workgroups (each CG launch re-reads that many partial sums; at
        // 128^3 ZU = 8 measured 7% faster than ZU = 4)
        zu_ = 16;
        for (int z : {4, 8, 16}) {
            if ((long)cols * ((nz + z - 1) / z) <= 1024) {
                zu_ = z;
                break;
            }
        }
        if (const char *e = std::getenv("AETHER_CG_ZU")) {
            const int z = std::atoi(e);
            if (z == 4 || z == 8 || z == 16) zu_ = z;
        }
        zu_default_ = zu_;
        tl_.gx = gx;
        tl_.gy = gy;
        tl_.gz = (nz + zu_ - 1) / zu_;
        tl_.nb = gx * gy * tl_.gz;
        tl_.remap = (tl_.nb % 8 == 0) ? 1 : 0;
        // two-launch CG with 16-byte lanes (four FP32 or two FP64 x-cells per
        // thread): 256-thread groups of txv lanes x ty rows x vzu planes (vzu
        // = 4 measured faster than 2 at 128^3 and 256^3 in FP32)
        if (nx % kLaneCells == 0) {
            const int w4 = nx / kLaneCells;
            tlv_.txv = std::min(w4, 64);
            tlv_.ty = kThreads / tlv_.txv;
            tlv_.gx = (w4 + tlv_.txv - 1) / tlv_.txv;
            tlv_.gy = (ny + tlv_.ty - 1) / tlv_.ty;
            vzu_ = 4;
            tlv_.gz = (nz + vzu_ - 1) / vzu_;
            tlv_.nb = tlv_.gx * tlv_.gy * tlv_.gz;
            tlv_.remap = (tlv_.nb % 8 == 0) ? 1 : 0;
        }
        // one-launch SOR sweep (k_rbsor4, FP32): the four-wide tiling with 2
        // planes per thread (0.79 ms per 128^3 RB-SOR step against 0.87 with 4)
        if (sizeof(T) == 4 && nx % 4 == 0) {
            tls_ = tlv_;
            tls_.gz = (nz + 1) / 2;
            tls_.nb = tls_.gx * tls_.gy * tls_.gz;
            tls_.remap = (tls_.nb % 8 == 0) ? 1 : 0;
        }
        // one-launch four-wide CG (cg_cgs4.hpp): the four-wide tiling with 4
        // planes per thread
        if (nx % 4 == 0) {
            const int w4 = nx / 4;
            tl4_.txv = std::min(w4, 64);
            tl4_.ty = kThreads / tl4_.txv;
            tl4_.gx = (w4 + tl4_.txv - 1) / tl4_.txv;
            tl4_.gy = (ny + tl4_.ty - 1) / tl4_.ty;
            tl4_.gz = (nz + 3) / 4;
            tl4_.nb = tl4_.gx * tl4_.gy * tl4_.gz;
            tl4_.remap = (tl4_.nb % 8 == 0) ? 1 : 0;
        }
        // one-launch CG with one cell per thread (nx % 4 != 0): 64 x 4 x 16 tiles
        tlf_.gx = gx;
        tlf_.gy = gy;
        tlf_.gz = (nz + kFusedKZ - 1) / kFusedKZ;
        tlf_.nb = gx * gy * tlf_.gz;
        tlf_.remap = (tlf_.nb % 8 == 0) ? 1 : 0;
        const int npart = std::max(std::max(std::max(nb_, tlv_.nb), std::max(tl_.nb, tlf_.nb)),
                                   tl4_.nb);
        // [0]: id of the last solve a CG launch stopped; [1..2]: one 64-bit
        // word, (id << 32) | iterations of the last finished solve.
        // Coherent, so device writes are seen without a cache flush.
        HIPT(hipHostMalloc((void **)&done_host_, 64,
                           hipHostMallocMapped | hipHostMallocCoherent));
        done_host_[0] = 0;
        done_host_[1] = 0;
        done_host_[2] = 0;
        HIPT(hipHostGetDevicePointer((void **)&done_dev_, done_host_, 0));
        for (int a = 0; a < 5; ++a) {
            f_[a].alloc(n_);
            alt_[a].alloc(n_);
            cur_[a] = f_[a].p;
            nxt_[a] = alt_[a].p;
        }
        fwd_[0].alloc(n_);
        fwd_[1].alloc(n_);
        curl_.alloc(n_);
        p_.alloc(n_);
        b_.alloc(n_);
        r_.alloc(n_);
        d_[0].alloc(n_);
        d_[1].alloc(n_);
        partA_.alloc(npart);
        partB_.alloc(npart);
        partC_.alloc(npart);
        partD64_.alloc(npart);
        tot_.alloc(4);
        part_tel_.alloc((size_t)nb_ * kTelParts);
        scalar_.alloc(8);
        tel_.alloc(8);
        cg_.alloc(1);
        iters_log_.alloc(1);
        nonfinite_.alloc(1);
        k_fill<int><<<1, kThreads, 0, st_>>>(nonfinite_.p, 1, 0);
        LAUNCH_CHECK();
        if (ir64_) {
            p64_.alloc(n_);
            b64_.alloc(n_);
            fill64(p64_.p);
            fill64(b64_.p);
        }
        for (int a = 0; a < 5; ++a) zero(cur_[a]);
        zero(p_.p);
        zero(b_.p);
        zero(r_.p);
        zero(d_[0].p);
        zero(d_[1].p);
        HIPT(hipMemsetAsync(cg_.p, 0, sizeof(CgState), st_));
        sync();
    }

    ~Solver() override {
        (void)hipStreamSynchronize(st_);
        for (auto e : chunk_events_) (void)hipEventDestroy(e);
        if (done_host_) (void)hipHostFree(done_host_);
    }

    void fill64(double *x) {
        k_fill<double><<<blocks_for(n_), kThreads, 0, st_>>>(x, (long)n_, 0.0);
        LAUNCH_CHECK();
    }

    void zero(T *x) {
        k_fill<T><<<blocks_for(n_), kThreads, 0, st_>>>(x, (long)n_, T(0));
        LAUNCH_CHECK();
    }
    // Workgroups of a grid-stride reduction writing part[blockIdx.x]: never
    // more than the partial-sum buffer holds.
    static unsigned partial_blocks(size_t n, const DevBuf<double> &part) {
        if (part.n == 0) raise(AETHER_EINTERNAL, "reduction into an unallocated buffer");
        return std::min<unsigned>(blocks_for(n), (unsigned)std::min<size_t>(part.n, 4096));
    }
    static unsigned blocks_for(size_t n) {
        size_t b = (n + kThreads - 1) / kThreads;
        return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 4096);
    }
    void sync() { HIPT(hipStreamSynchronize(st_)); }

    // --- stages ----------------------------------------------------------
    void advect(int kind, double dt, double diss) override {
        if (kind < 0 || kind > 4) raise(AETHER_EINVAL, "field_kind must be 0..4");
        enqueue_advect1(kind, dt, diss);
        sync();
    }
    void enqueue_advect1(int kind, double dt, double diss) {
        const T s = (T)(dt / h_);  // src/fluid.rs:154
        T *src = cur_[kind];
        k_sl_forward<T, 1><<<cell_grid_, cell_blk_, 0, st_>>>(g_, s, cur_[0], cur_[1], cur_[2],
                                                              src, nullptr, fwd_[0].p, nullptr);
        LAUNCH_CHECK();
        k_mac_correct<T, 1><<<cell_grid_, cell_blk_, 0, st_>>>(
            g_, s, cur_[0], cur_[1], cur_[2], src, fwd_[0].p, nxt_[kind], (T)diss, nullptr,
            nullptr, nullptr, T(0), kind >= 3 ? 1 : 0);
        LAUNCH_CHECK();
        std::swap(cur_[kind], nxt_[kind]);
    }
    // The three in-place velocity advections of a step (src/fluid.rs:599-601)
    // in four passes: each component's correction also traces the next
    // component along the velocity it leaves (k_mac_correct_forward).  The
    // forward fields alternate between the two buffers, since a correction
    // gathers its own forward field around each cell while it writes the next.
    void enqueue_advect_velocity(double dt, double diss) {
        const T s = (T)(dt / h_);  // src/fluid.rs:154
        k_sl_forward<T, 1><<<cell_grid_, cell_blk_, 0, st_>>>(g_, s, cur_[0], cur_[1], cur_[2],
                                                              cur_[0], nullptr, fwd_[0].p,
                                                              nullptr);
        LAUNCH_CHECK();
        for (int kind = 0; kind < 2; ++kind) {
            k_mac_correct_forward<T><<<cell_grid_, cell_blk_, 0, st_>>>(
                g_, s, cur_[0], cur_[1], cur_[2], cur_[kind], fwd_[kind].p, nxt_[kind], (T)diss,
                kind, cur_[kind + 1], fwd_[kind ^ 1].p);
            LAUNCH_CHECK();
            std::swap(cur_[kind], nxt_[kind]);
        }
        k_mac_correct<T, 1><<<cell_grid_, cell_blk_, 0, st_>>>(
            g_, s, cur_[0], cur_[1], cur_[2], cur_[2], fwd_[0].p, nxt_[2], (T)diss, nullptr,
            nullptr, nullptr, T(0), 0);
        LAUNCH_CHECK();
        std::swap(cur_[2], nxt_[2]);
    }
    // density and temperature share one velocity: one fused pass pair.
    // with_gradient: the projection left its gradient subtraction to this
    // pass, which forms the projected velocity and traces along it at once.
    void enqueue_advect_scalars(double dt, bool with_gradient = false) {
        const T s = (T)(dt / h_);
        if (with_gradient)
            k_gradient_sl_forward<T><<<cell_grid_, cell_blk_, 0, st_>>>(
                g_, (T)(1.0 / h_), p_.p, cur_[0], cur_[1], cur_[2], s, cur_[3], cur_[4],
                fwd_[0].p, fwd_[1].p);
        else
            k_sl_forward<T, 2><<<cell_grid_, cell_blk_, 0, st_>>>(
                g_, s, cur_[0], cur_[1], cur_[2], cur_[3], cur_[4], fwd_[0].p, fwd_[1].p);
        LAUNCH_CHECK();
        k_mac_correct<T, 2><<<cell_grid_, cell_blk_, 0, st_>>>(
            g_, s, cur_[0], cur_[1], cur_[2], cur_[3], fwd_[0].p, nxt_[3],
            (T)params.dissipation_density, cur_[4], fwd_[1].p, nxt_[4],
            (T)params.dissipation_temperature, 1);
        LAUNCH_CHECK();
        std::swap(cur_[3], nxt_[3]);
        std::swap(cur_[4], nxt_[4]);
    }

    void enforce_bc() override {
        k_enforce_bc<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, cur_[0], cur_[1], cur_[2]);
        LAUNCH_CHECK();
        sync();
    }

    void buoyancy_confinement(double dt) override {
        enqueue_buoyancy_confinement(dt);
        sync();
    }
    // with_sphere: also apply the sphere source, which the step runs just
    // before (the curl pass reads no density or temperature)
    void enqueue_buoyancy_confinement(double dt, bool with_sphere = false) {
        const double inv_2h = 0.5 / h_;                    // src/fluid.rs:257
        const double eps_h = params.vorticity_eps * h_;   // src/fluid.rs:284
        SphereSrc sph = {0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (with_sphere && params.source_radius > 0.0) {
            const double r = params.source_radius;
            sph = {1,     params.source_center[0], params.source_center[1], params.source_center[2],
                   r * r, params.source_density,   params.source_temperature};
        }
        k_curl_mag<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, (T)inv_2h, cur_[0], cur_[1],
                                                          cur_[2], curl_.p, sph, cur_[3], cur_[4]);
        LAUNCH_CHECK();
        k_confine_buoy_bc<T><<<cell_grid_, cell_blk_, 0, st_>>>(
            g_, (T)dt, (T)inv_2h, (T)eps_h, (T)params.buoyancy_beta, (T)params.gravity_alpha,
            cur_[0], cur_[1], cur_[2], curl_.p, cur_[3], cur_[4], nxt_[0], nxt_[1], nxt_[2]);
        LAUNCH_CHECK();
        for (int a = 0; a < 3; ++a) std::swap(cur_[a], nxt_[a]);
    }

    double read_scalar() {
        double v = 0;
        HIPT(hipMemcpyAsync(&v, scalar_.p, sizeof(double), hipMemcpyDeviceToHost, st_));
        sync();
        return v;
    }

    double divergence_norm() override {
        // src/fluid.rs:333-353: stores b and returns the interior RMS of div u
        k_divergence<T><<<red_grid_, cell_blk_, 0, st_>>>(g_, zchunk_, (T)h_, (T)(1.0 / h_), 0,
                                                           cur_[0], cur_[1], cur_[2], b_.p,
                                                           partA_.p);
        LAUNCH_CHECK();
        k_partials_final<<<1, kThreads, 0, st_>>>(partA_.p, nb_, 0, scalar_.p);
        LAUNCH_CHECK();
        const double sum_sq = read_scalar();
        const double interior_n = (double)((size_t)(g_.nx - 2) * (g_.ny - 2) * (g_.nz - 2));
        return std::sqrt(sum_sq / interior_n);
    }

    double divergence_max() override {
        k_div_max<T><<<red_grid_, cell_blk_, 0, st_>>>(g_, zchunk_, (T)(1.0 / h_), cur_[0],
                                                        cur_[1], cur_[2], partA_.p);
        LAUNCH_CHECK();
        k_partials_final<<<1, kThreads, 0, st_>>>(partA_.p, nb_, 1, scalar_.p);
        LAUNCH_CHECK();
        return read_scalar();
    }

    double rel_tol2() const {
        const double t = rel_override_ >= 0.0 ? rel_override_ : params.cg_rel_tol;
        return t > 0.0 ? t * t : 0.0;
    }

    // CG iterations (src/fluid.rs:463-487) as two launches each; every
    // launch returns at once after the device stop flag is set.
    // Launch with start/stop events stamped by the dispatch itself (kernel
    // execution only, no queueing gap): the kernel-timing mode.  Every
    // argument must have exactly the kernel parameter's type.
    template <typename... Args>
    void launch_stamped(const void *fn, hipEvent_t e0, hipEvent_t e1, Args... args) {
        void *ptrs[] = {(void *)&args...};
        HIPT(hipExtLaunchKernel(fn, dim3(tl_.nb), cell_blk_, ptrs, 0, st_, e0, e1, 0));
    }

    // One-launch CG iteration `it` (cg_fused.hpp).  gamma partials alternate
    // between partB_ (even it: written by the initialisation) and partA_,
    // delta partials between partE_ and partF_, r between r_ and r2_.
    int fused_nb() const { return tl4_.nb > 0 ? tl4_.nb : tlf_.nb; }
    // write-through stores address an FP32 array through 32-bit byte offsets
    bool wt_ok() const { return n_ * 4 < (size_t(1) << 31); }

    void enqueue_cg_fused(int it, double rel2, int solve_id, int step_index) {
        const int nb_in = it == 0 ? tl_.nb : fused_nb();
        const double *pg_in = (it & 1) ? partA_.p : partB_.p;
        const double *pd_in = (it & 1) ? partF_.p : partE_.p;
        double *pg_out = (it & 1) ? partB_.p : partA_.p;
        double *pd_out = (it & 1) ? partE_.p : partF_.p;
        const double *pbb = partC_.p;
        const T *r_in = (it & 1) ? r2_.p : r_.p;
        T *r_out = (it & 1) ? r_.p : r2_.p;
        const T *d_in = d_[(it + 1) & 1].p;
        T *d_out = d_[it & 1].p;
        T *pp = p_.p;
        CgState *st = cg_.p;
        int *dm = done_dev_;
        if constexpr (sizeof(T) == 4) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timer.enabled) {
                e0 = timer.get();
                e1 = timer.get();
                timer.recs.push_back({3, step_index, it, e0, e1});
            }
            if (tl4_.nb > 0) {  // four-wide, A r carried (cg_cgs4.hpp)
                Dims g = g_;
                VTiling vt = tl4_;
                int itv = it, nbv = nb_in, sid = solve_id;
                double rt = rel2;
                const T *w_in = w_[it & 1].p;
                T *w_out = w_[(it + 1) & 1].p, *dd = d_[0].p;
                // r rotates through a ring of kCgsCycle arrays: launch `it` reads
                // ring[it % K], writes ring[(it + 1) % K] (which holds
                // r_(it+1-K)) and finds r_(it-1) in ring[(it - 1) % K]
                CgRing ring{};
                for (int q = 0; q < kCgsCycle; ++q) ring.r[q] = ring_at(q);
                void *args[] = {&g,      &vt,     &itv, &rt,   &nbv,  &pg_in, &pd_in, &pbb, &pg_out,
                                &pd_out, &st,     &ring, &w_in, &w_out, &dd,   &pp,    &dm,  &sid};
                // write-through stores for arrays under 2 GiB
                const int rg = !vtiling_regular(vt, g.ny, g.nz, 4) ? 0 : vt.txv == 32 ? 2 : 1;
                const void *fn = cgs4_kernel(it == 0, it % kCgsCycle, std::max(nb_in, vt.nb),
                                             wt_ok(), rg);
                HIPT(hipExtLaunchKernel(fn, dim3(vt.nb), dim3(kThreads), args, 0, st_, e0, e1, 0));
                return;
            }
            // one cell per thread (nx % 4 != 0): radius-2 recompute (cg_fused.hpp)
            const void *fn = it == 0 ? (const void *)k_cg_fused<T, kFusedKZ, kFusedPF, true>
                                     : (const void *)k_cg_fused<T, kFusedKZ, kFusedPF, false>;
            Dims g = g_;
            Tiling tl = tlf_;
            int itv = it, nbv = nb_in, sid = solve_id;
            double rt = rel2;
            void *args[] = {&g,      &tl, &itv,  &rt,   &nbv,  &pg_in, &pd_in, &pbb, &pg_out,
                            &pd_out, &st, &r_in, &r_out, &d_in, &d_out, &pp,   &dm,  &sid};
            HIPT(hipExtLaunchKernel(fn, dim3(tlf_.nb), cell_blk_, args, 0, st_, e0, e1, 0));
        } else {
            raise(AETHER_EINVAL, "the one-launch CG is built for FP32 only");
        }
    }

    // The CG layout of the next solve (AETHER_CG_SCALAR/X4/FUSED).
    // Untuned choice for FP32.  The four-wide one-launch CG moves 8 array
    // passes per iteration on average in one launch, the two-launch form 7.5
    // in two: the one launch wins while its eight FP32 arrays (32 N bytes) stay resident
    // in the 256 MB Infinity Cache next to the step's other fields (measured:
    // 20% faster at 96^3, 16% at 128^3, slower from 192^3 on).
    int default_form() const {
        if (g_.nx % 4 != 0) return AETHER_CG_SCALAR;
        return n_ <= kOneLaunchMaxCells ? AETHER_CG_FUSED : AETHER_CG_X4;
    }
    static constexpr size_t kOneLaunchMaxCells = 3200000;  // 32 N bytes <= ~100 MB

    int effective_form() const {
        if (trial_form_) return trial_form_;  // the autotuner's candidate
        int f = params.cg_form;
        if constexpr (sizeof(T) != 4) {
            // FP64: two cells per 16-byte lane when nx % 2 == 0 (measured 1.38x
            // faster per iteration than the best one-cell-per-thread tiling at
            // 128^3, 1.89x at 256^3); the one-launch form is built for FP32
            if (f == AETHER_CG_AUTO)
                f = tuned_form_ ? tuned_form_ : tlv_.nb > 0 ? AETHER_CG_X4 : AETHER_CG_SCALAR;
            if (f == AETHER_CG_FUSED) f = AETHER_CG_SCALAR;
            if (f == AETHER_CG_X4 && tlv_.nb == 0)
                raise(AETHER_EINVAL, "AETHER_CG_X4 in FP64 needs nx % 2 == 0");
        } else {
            if (f == AETHER_CG_AUTO)
                f = tuned_form_ ? tuned_form_ : default_form();
            if (f == AETHER_CG_X4 && g_.nx % 4 != 0)
                raise(AETHER_EINVAL, "AETHER_CG_X4 needs nx % 4 == 0");
        }
        return f;
    }

    // Partial-sum buffers hold at least n slots (their contents are per solve).
    void ensure_partials(int n) {
        for (DevBuf<double> *b : {&partA_, &partB_, &partC_})
            if ((int)b->n < n) b->alloc(n);
        if (r2_.p)
            for (DevBuf<double> *b : {&partE_, &partF_})
                if ((int)b->n < n) b->alloc(n);
    }

    // z cells per thread of the one-cell-per-thread tiling (also the CG
    // initialisation's tiling) and of the four-wide tiling.
    void set_tilings(int zu, int vzu) {
        zu_ = zu;
        tl_.gz = (g_.nz + zu_ - 1) / zu_;
        tl_.nb = tl_.gx * tl_.gy * tl_.gz;
        tl_.remap = (tl_.nb % 8 == 0) ? 1 : 0;
        ensure_partials(tl_.nb);
        if (tlv_.nb > 0) {
            vzu_ = vzu;
            tlv_.gz = (g_.nz + vzu_ - 1) / vzu_;
            tlv_.nb = tlv_.gx * tlv_.gy * tlv_.gz;
            tlv_.remap = (tlv_.nb % 8 == 0) ? 1 : 0;
            ensure_partials(tlv_.nb);
        }
    }

    // Tilings of the next solve: cg_z_cells when it is set (the four-wide
    // layout takes 2 or 4, the one-cell-per-thread layout 4, 8 or 16), else
    // the autotuner's choice, else the defaults chosen at creation.
    void sync_tilings() {
        const int zc = params.cg_z_cells;
        int zu = tuned_zu_ ? tuned_zu_ : zu_default_;
        int vzu = tuned_vzu_ ? tuned_vzu_ : 4;
        if (zc == 2 || zc == 4) vzu = zc;
        if (zc == 4 || zc == 8 || zc == 16) zu = zc;
        if (zu != zu_ || (tlv_.nb > 0 && vzu != vzu_)) set_tilings(zu, vzu);
    }

    // Times `iters` CG iterations of every candidate layout on a synthetic
    // right-hand side of this grid (p and b are saved and restored) and
    // keeps the fastest as this handle's AETHER_CG_AUTO choice.  us[6]
    // receives the time per iteration of: four-wide with 2 and 4 z cells per
    // thread, one cell per thread with 4, 8 and 16, the one-launch form
    // (-1 where a candidate does not apply).
    void autotune(int iters, double *us, int *form_out, int *z_out) override {
        if (iters < 1) raise(AETHER_EINVAL, "iters must be >= 1");
        struct Cand {
            int form, z;
        };
        const Cand cands[6] = {{AETHER_CG_X4, 2},     {AETHER_CG_X4, 4},
                               {AETHER_CG_SCALAR, 4}, {AETHER_CG_SCALAR, 8},
                               {AETHER_CG_SCALAR, 16}, {AETHER_CG_FUSED, 4}};
        const bool fp32 = sizeof(T) == 4;
        DevBuf<T> saved_p(n_), saved_b(n_);
        HIPT(hipMemcpyAsync(saved_p.p, p_.p, n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
        HIPT(hipMemcpyAsync(saved_b.p, b_.p, n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
        k_fill_hash<T><<<blocks_for(n_), kThreads, 0, st_>>>(b_.p, (long)n_);
        LAUNCH_CHECK();
        const int zu0 = zu_, vzu0 = vzu_;
        const double rel0 = rel_override_;
        rel_override_ = 0.0;  // absolute guards only: every candidate runs `iters`
        hipEvent_t e0, e1;
        HIPT(hipEventCreate(&e0));
        HIPT(hipEventCreate(&e1));
        double best = 1e300;
        Cand pick = {AETHER_CG_SCALAR, zu0};
        try {
            for (int a = 0; a < 6; ++a) {
                const Cand c = cands[a];
                const bool ok = c.form == AETHER_CG_SCALAR ||
                                (c.form == AETHER_CG_X4 && tlv_.nb > 0) ||
                                (fp32 && c.form == AETHER_CG_FUSED);
                if (us) us[a] = -1.0;
                if (!ok) continue;
                set_tilings(c.form == AETHER_CG_SCALAR ? c.z : zu0,
                            c.form == AETHER_CG_X4 ? c.z : vzu0);
                trial_form_ = c.form;
                float ms = 0.0f;
                for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
                    zero(p_.p);
                    ZU_DISPATCH(zu_, (k_div_cginit<T, ZU><<<tl_.nb, cell_blk_, 0, st_>>>(
                                          g_, tl_, (T)h_, 0, cur_[0], cur_[1], cur_[2], p_.p,
                                          b_.p, r_.p, partB_.p, partC_.p, cg_.p)));
                    LAUNCH_CHECK();
                    HIPT(hipEventRecord(e0, st_));
                    enqueue_cg_loop(iters, -1);
                    HIPT(hipEventRecord(e1, st_));
                    HIPT(hipEventSynchronize(e1));
                    HIPT(hipEventElapsedTime(&ms, e0, e1));
                }
                const double t = 1e3 * (double)ms / iters;
                if (us) us[a] = t;
                if (t < best) {
                    best = t;
                    pick = c;
                }
            }
        } catch (...) {
            trial_form_ = 0;
            rel_override_ = rel0;
            set_tilings(zu0, vzu0);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        trial_form_ = 0;
        rel_override_ = rel0;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        set_tilings(pick.form == AETHER_CG_SCALAR ? pick.z : zu0,
                    pick.form == AETHER_CG_X4 ? pick.z : vzu0);
        tuned_form_ = pick.form;
        tuned_zu_ = zu_;
        tuned_vzu_ = vzu_;
        HIPT(hipMemcpyAsync(p_.p, saved_p.p, n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
        HIPT(hipMemcpyAsync(b_.p, saved_b.p, n_ * sizeof(T), hipMemcpyDeviceToDevice, st_));
        sync();
        if (form_out) *form_out = pick.form;
        if (z_out) *z_out = pick.z;
    }

    void ensure_fused_buffers() {
        if (!r2_.p) {
            const size_t npart = (size_t)std::max(
                std::max(std::max(nb_, tlv_.nb), std::max(tl_.nb, tlf_.nb)), tl4_.nb);
            r2_.alloc(n_);
            zero(r2_.p);
            partE_.alloc(npart);
            partF_.alloc(npart);
        }
        if (sizeof(T) == 4 && tl4_.nb > 0) {
            if (!w_[0].p)
                for (DevBuf<T> *b : {&w_[0], &w_[1]}) {
                    b->alloc(n_);
                    zero(b->p);
                }
            for (int q = 0; q + 2 < kCgsCycle; ++q)  // the residual ring past r_ and r2_
                if (!rring_[q].p) {
                    rring_[q].alloc(n_);
                    zero(rring_[q].p);
                }
        }
    }

    // Residual array q of the one-launch four-wide CG's ring.
    T *ring_at(int q) { return q == 0 ? r_.p : q == 1 ? r2_.p : rring_[q - 2].p; }

    // Four-wide two-launch CG iteration `it` (cg_vec.hpp); same buffers and
    // partial-sum roles as the one-cell-per-thread kernels.
    void enqueue_cg_x4(int it, double rel2, int solve_id, int step_index) {
        if constexpr (sizeof(T) == 4) {
            const int nb_in = it == 0 ? tl_.nb : tlv_.nb;
            hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
            if (timer.enabled) {
                e0 = timer.get();
                e1 = timer.get();
                e2 = timer.get();
                e3 = timer.get();
                timer.recs.push_back({0, step_index, it, e0, e1});
                timer.recs.push_back({1, step_index, it, e2, e3});
            }
            Dims g = g_;
            VTiling vt = tlv_;
            int itv = it, nbv = nb_in, sid = solve_id;
            double rt = rel2;
            const double *prr = partB_.p, *pbb = partC_.p, *pdq_in = partA_.p;
            double *pdq = partA_.p, *prr_out = partB_.p;
            CgState *st = cg_.p;
            const T *r = r_.p, *dold = d_[(it + 1) & 1].p, *dcur = d_[it & 1].p;
            T *dnew = d_[it & 1].p, *pp = p_.p, *rw = r_.p;
            int *dm = done_dev_;
            // Past 1024 partials, the direction kernel's workgroups re-reading
            // all of them costs more than one extra launch (at 256^3, 4096 x
            // 32 KB of L2 reads: ~12 us against ~3 us), so one workgroup sums
            // them first, in the order the kernel would use, and the
            // direction kernel reads one total.  The update kernel hides its
            // partial loads behind its five passes and keeps summing them.
            const bool wide = std::max(nb_in, vt.nb) > 4 * kThreads;  // > 1024 partials
            int nbu = vt.nb;
            if (wide) {
                k_partials_total<<<1, kThreads, 0, st_>>>(prr, nb_in, tot_.p);
                if (it == 0) k_partials_total<<<1, kThreads, 0, st_>>>(pbb, nb_in, tot_.p + 1);
                LAUNCH_CHECK();
                prr = tot_.p;
                pbb = tot_.p + 1;
                nbv = 1;
            }
            const bool wt = wt_ok();  // write-through stores (see Out4)
            // p is brought up to date on odd iterations (body_cg_update4)
            const bool odd = it & 1;
            const T *dprev = dold;  // d_(it-1), for the odd update
#define X4_PICK(NQD, NQU, ST_)                                                             \
    do {                                                                                   \
        f1 = it == 0 ? (const void *)k_cg_dir4<VZU, true, NQD, ST_, false>                 \
             : odd   ? (const void *)k_cg_dir4<VZU, false, NQD, ST_, true>                 \
                     : (const void *)k_cg_dir4<VZU, false, NQD, ST_, false>;               \
        f2 = odd ? (const void *)k_cg_update4<VZU, NQU, ST_, kPBoth>                       \
                 : (const void *)k_cg_update4<VZU, NQU, ST_, kPSkip>;                      \
    } while (0)
            VZU_DISPATCH(vzu_, ({
                const void *f1, *f2;
                if (wide && wt)
                    X4_PICK(1, 16, kStWT);
                else if (wide)
                    X4_PICK(1, 16, kStPlain);
                else if (wt)
                    X4_PICK(4, 4, kStWT);
                else
                    X4_PICK(4, 4, kStPlain);
                void *a1[] = {&g, &vt, &itv, &rt, &nbv, &prr, &pbb, &st, &r, &dold, &dnew, &pdq,
                              &dm, &sid, &pp};
                HIPT(hipExtLaunchKernel(f1, dim3(vt.nb), dim3(kThreads), a1, 0, st_, e0, e1, 0));
                void *a2[] = {&g, &vt, &itv, &nbu, &pdq_in, &st, &dcur, &pp, &rw, &prr_out, &dm,
                              &sid, &dprev};
                HIPT(hipExtLaunchKernel(f2, dim3(vt.nb), dim3(kThreads), a2, 0, st_, e2, e3, 0));
            }));
#undef X4_PICK
        } else {
            // FP64: two x-cells per thread (cg_vec2.hpp), the same buffer and
            // partial-sum roles and the same p deferral as the FP32 kernels
            const int nb_in = it == 0 ? tl_.nb : tlv_.nb;
            hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
            if (timer.enabled) {
                e0 = timer.get();
                e1 = timer.get();
                e2 = timer.get();
                e3 = timer.get();
                timer.recs.push_back({0, step_index, it, e0, e1});
                timer.recs.push_back({1, step_index, it, e2, e3});
            }
            Dims g = g_;
            VTiling vt = tlv_;
            int itv = it, nbv = nb_in, sid = solve_id;
            double rt = rel2;
            const double *prr = partB_.p, *pbb = partC_.p, *pdq_in = partA_.p;
            double *pdq = partA_.p, *prr_out = partB_.p;
            CgState *st = cg_.p;
            const T *r = r_.p, *dold = d_[(it + 1) & 1].p, *dcur = d_[it & 1].p;
            T *dnew = d_[it & 1].p, *pp = p_.p, *rw = r_.p;
            int *dm = done_dev_;
            // Past 1024 partials one workgroup sums the previous launch's
            // first, as in the FP32 form; past 4096 the d.q partials too (two
            // cells per thread make twice the FP32 form's workgroups: 8192 at
            // 256^3 with 4 planes per thread).
            const bool wide = std::max(nb_in, vt.nb) > 4 * kThreads;
            const bool wide_u = vt.nb > 16 * kThreads;
            int nbu = vt.nb;
            if (wide) {
                k_partials_total<<<1, kThreads, 0, st_>>>(prr, nb_in, tot_.p);
                if (it == 0) k_partials_total<<<1, kThreads, 0, st_>>>(pbb, nb_in, tot_.p + 1);
                LAUNCH_CHECK();
                prr = tot_.p;
                pbb = tot_.p + 1;
                nbv = 1;
            }
            const bool wt = n_ * sizeof(T) < (size_t(1) << 31);  // 32-bit byte offsets
            const bool odd = it & 1;
            const T *dprev = dold;  // d_(it-1), for the odd update
#define X2_PICK(NQD, ST_)                                                                  \
    do {                                                                                   \
        f1 = it == 0 ? (const void *)k_cg_dir2<VZU, true, NQD, ST_, false>                 \
             : odd   ? (const void *)k_cg_dir2<VZU, false, NQD, ST_, true>                 \
                     : (const void *)k_cg_dir2<VZU, false, NQD, ST_, false>;               \
        if (nbu > 4 * kThreads)                                                            \
            f2 = odd ? (const void *)k_cg_update2<VZU, 16, ST_, kPBoth>                    \
                     : (const void *)k_cg_update2<VZU, 16, ST_, kPSkip>;                   \
        else                                                                               \
            f2 = odd ? (const void *)k_cg_update2<VZU, 4, ST_, kPBoth>                     \
                     : (const void *)k_cg_update2<VZU, 4, ST_, kPSkip>;                    \
    } while (0)
            VZU_DISPATCH(vzu_, ({
                const void *f1, *f2;
                if (wide_u) nbu = 1;  // summed below, after the direction launch
                if (wide && wt)
                    X2_PICK(1, kStWT);
                else if (wide)
                    X2_PICK(1, kStPlain);
                else if (wt)
                    X2_PICK(4, kStWT);
                else
                    X2_PICK(4, kStPlain);
                void *a1[] = {&g, &vt, &itv, &rt, &nbv, &prr, &pbb, &st, &r, &dold, &dnew, &pdq,
                              &dm, &sid, &pp};
                HIPT(hipExtLaunchKernel(f1, dim3(vt.nb), dim3(kThreads), a1, 0, st_, e0, e1, 0));
                if (wide_u) {
                    k_partials_total<<<1, kThreads, 0, st_>>>(pdq, vt.nb, tot_.p + 2);
                    LAUNCH_CHECK();
                    pdq_in = tot_.p + 2;
                }
                void *a2[] = {&g, &vt, &itv, &nbu, &pdq_in, &st, &dcur, &pp, &rw, &prr_out, &dm,
                              &sid, &dprev};
                HIPT(hipExtLaunchKernel(f2, dim3(vt.nb), dim3(kThreads), a2, 0, st_, e2, e3, 0));
            }));
#undef X2_PICK
        }
    }

    void enqueue_cg_iteration(int it, double rel2, int solve_id, int step_index) {
        if (form_ == AETHER_CG_FUSED) {
            enqueue_cg_fused(it, rel2, solve_id, step_index);
            return;
        }
        if (form_ == AETHER_CG_X4) {
            enqueue_cg_x4(it, rel2, solve_id, step_index);
            return;
        }
        const double *pB = partB_.p, *pC = partC_.p, *pA = partA_.p;
        const T *dold = d_[(it + 1) & 1].p, *dcur = d_[it & 1].p, *rr = r_.p;
        if (timer.enabled) {
            hipEvent_t e0 = timer.get(), e1 = timer.get(), e2 = timer.get(), e3 = timer.get();
            ZU_DISPATCH(zu_, (launch_stamped(
                                 it == 0 ? (const void *)k_cg_dir<T, ZU, true>
                                         : (const void *)k_cg_dir<T, ZU, false>,
                                 e0, e1, g_, tl_, it, rel2, pB, pC, cg_.p, rr, dold,
                                 d_[it & 1].p, partA_.p, done_dev_, solve_id)));
            ZU_DISPATCH(zu_, (launch_stamped((const void *)k_cg_update<T, ZU>, e2, e3, g_, tl_,
                                             it, pA, cg_.p, dcur, p_.p, r_.p, partB_.p,
                                             done_dev_, solve_id)));
            timer.recs.push_back({0, step_index, it, e0, e1});
            timer.recs.push_back({1, step_index, it, e2, e3});
            return;
        }
        if (it == 0)
            ZU_DISPATCH(zu_, (k_cg_dir<T, ZU, true><<<tl_.nb, cell_blk_, 0, st_>>>(
                                  g_, tl_, it, rel2, partB_.p, partC_.p, cg_.p, r_.p,
                                  d_[(it + 1) & 1].p, d_[it & 1].p, partA_.p, done_dev_,
                                  solve_id)));
        else
            ZU_DISPATCH(zu_, (k_cg_dir<T, ZU, false><<<tl_.nb, cell_blk_, 0, st_>>>(
                                  g_, tl_, it, rel2, partB_.p, partC_.p, cg_.p, r_.p,
                                  d_[(it + 1) & 1].p, d_[it & 1].p, partA_.p, done_dev_,
                                  solve_id)));
        LAUNCH_CHECK();
        ZU_DISPATCH(zu_, (k_cg_update<T, ZU><<<tl_.nb, cell_blk_, 0, st_>>>(
                              g_, tl_, it, partA_.p, cg_.p, d_[it & 1].p, p_.p, r_.p, partB_.p,
                              done_dev_, solve_id)));
        LAUNCH_CHECK();
    }

    // CG iterations (src/fluid.rs:463-487).  Every scalar stays on the
    // device; the host only watches a mapped "solve finished" word, two
    // chunks behind, to stop queueing launches once the device has stopped
    // (launches queued after the stop return without touching memory other
    // than their loads).  When one of the last two solves finished at the
    // cap, the whole solve is queued at once: runs of capped solves (the
    // 128^3 smoke step) then wait on no event.  Solves that stop early keep
    // the paced chunks, since a long first chunk would queue early-exit
    // launches.  The host runs up to two solves ahead of the device, so the
    // latest finished solve may be the one before the previous; an older
    // count is not used.  After a capped solve, at most two whole solves of
    // early-exit launches are queued before a converging one is seen.
    void enqueue_cg_loop(int max_iters, int step_index) {
        const double rel2 = rel_tol2();
        const int solve_id = ++solve_seq_;
        const int chunk = kCgChunk;
        int first = chunk;
        {
            const unsigned long long last = __atomic_load_n(
                reinterpret_cast<unsigned long long *>(done_host_ + 2), __ATOMIC_ACQUIRE);
            const int last_id = (int)(last >> 32), last_iters = (int)(last & 0xffffffffu);
            if (last_id > 0 && last_id >= solve_id - 2 && last_iters >= max_iters)
                first = std::max(chunk, max_iters);
        }
        // chunk c: launches [c_begin(c), c_begin(c + 1)) (capped at max_iters)
        auto c_begin = [&](int c) { return c == 0 ? 0 : first + (c - 1) * chunk; };
        const int nchunks = max_iters > first ? 1 + (max_iters - first + chunk - 1) / chunk : 1;
        if ((int)chunk_events_.size() < nchunks) {
            for (int a = (int)chunk_events_.size(); a < nchunks; ++a) {
                hipEvent_t e;
                // pacing only: the done flag they gate is written straight to
                // coherent host memory, so no system-scope cache fence is
                // needed (with one, every record idled the GPU ~6 us)
                HIPT(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
                chunk_events_.push_back(e);
            }
        }
        form_ = effective_form();
        if (form_ == AETHER_CG_FUSED) ensure_fused_buffers();
        if (form_ == AETHER_CG_FUSED && max_iters > 0) {  // delta_0 = r_0 . A r_0 partials
            T *w0 = (sizeof(T) == 4 && tl4_.nb > 0) ? w_[0].p : nullptr;  // w_0 = A r_0
            ZU_DISPATCH(zu_, (k_cg_rar<T, ZU><<<tl_.nb, cell_blk_, 0, st_>>>(g_, tl_, r_.p,
                                                                             partE_.p, w0)));
            LAUNCH_CHECK();
        }
        for (int c = 0; c < nchunks; ++c) {
            if (c >= 2) {
                HIPT(hipEventSynchronize(chunk_events_[c - 2]));
                if (__atomic_load_n(done_host_, __ATOMIC_ACQUIRE) == solve_id) break;
            }
            const int it_end = std::min(max_iters, c_begin(c + 1));
            for (int it = c_begin(c); it < it_end; ++it)
                enqueue_cg_iteration(it, rel2, solve_id, step_index);
            HIPT(hipEventRecord(chunk_events_[c], st_));
        }
        if constexpr (sizeof(T) == 4) {
            // the four-wide two-launch form brings p up to date on odd
            // iterations, the one-launch form once every kCgsCycle launches: a
            // cap that ends the solve on another iteration leaves terms pending
            const bool fused4 = form_ == AETHER_CG_FUSED && tl4_.nb > 0;
            if (max_iters > 0 && ((fused4 && max_iters % kCgsCycle != 0) ||
                                  (form_ == AETHER_CG_X4 && (max_iters & 1)))) {
                const long n4 = (long)(n_ / 4);
                const unsigned nbt =
                    (unsigned)std::min<long>((n4 + kThreads - 1) / kThreads, 4096);
                const int last = max_iters - 1;
                if (fused4) {  // the one-launch form also left the directions unbuilt
                    // the residuals the pending launches read, in launch order
                    // (launch `it` read ring[it % K])
                    CgRing pend{};
                    const int j0 = last - last % kCgsCycle;
                    for (int q = 0; q <= last % kCgsCycle; ++q)
                        pend.r[q] = ring_at((j0 + q) % kCgsCycle);
                    cgs4_tail_launch(nbt, st_, cg_.p, last, pend,
                                     reinterpret_cast<const float4 *>(d_[0].p),
                                     reinterpret_cast<float4 *>(p_.p), n4);
                } else {
                    k_cg_p_tail<<<nbt, kThreads, 0, st_>>>(
                        cg_.p, last, reinterpret_cast<const float4 *>(d_[last & 1].p),
                        reinterpret_cast<float4 *>(p_.p), n4);
                }
                LAUNCH_CHECK();
            }
        } else if (form_ == AETHER_CG_X4 && max_iters > 0 && (max_iters & 1)) {
            // the FP64 two-wide form defers p as the FP32 one
            const long n2 = (long)(n_ / 2);
            const unsigned nbt = (unsigned)std::min<long>((n2 + kThreads - 1) / kThreads, 4096);
            const int last = max_iters - 1;
            k_cg_p_tail2<<<nbt, kThreads, 0, st_>>>(
                cg_.p, last, reinterpret_cast<const double2 *>(d_[last & 1].p),
                reinterpret_cast<double2 *>(p_.p), n2);
            LAUNCH_CHECK();
        }
        const double *part_rr = partB_.p;
        int nb_rr = tl_.nb;
        if (form_ == AETHER_CG_FUSED && max_iters > 0) {  // gamma of the last launch
            part_rr = (max_iters & 1) ? partA_.p : partB_.p;
            nb_rr = fused_nb();
        } else if (form_ == AETHER_CG_X4 && max_iters > 0) {
            nb_rr = tlv_.nb;
        }
        k_cg_finalize<<<1, cell_blk_, 0, st_>>>(max_iters, rel2, part_rr, partC_.p, nb_rr, cg_.p,
                                                 step_index >= 0 ? iters_dev_.p + step_index
                                                                 : iters_log_.p,
                                                 reinterpret_cast<unsigned long long *>(
                                                     done_dev_ + 2),
                                                 solve_id);
        LAUNCH_CHECK();
        last_kind_ = AETHER_SOLVER_CG;
    }

    void enqueue_rbgs(int sweeps) {
        bool four = false;
        if constexpr (sizeof(T) == 4) four = tlv_.nb > 0 && params.cg_form != AETHER_CG_SCALAR;
        if (four) {
            if constexpr (sizeof(T) == 4) {
                // one launch per sweep when every row fills whole wave
                // segments (k_rbsor4), from p to a second array and back
                if (tls_.gx == 1 && 64 % tls_.txv == 0 && sweeps > 0) {
                    if (!p2_.p) p2_.alloc(n_);
                    for (int it = 0; it < sweeps; ++it) {
                        const T *pin = (it & 1) ? p2_.p : p_.p;
                        T *pout = (it & 1) ? p_.p : p2_.p;
                        if (wt_ok())
                            k_rbsor4<2, kStWT><<<tls_.nb, kThreads, 0, st_>>>(
                                g_, tls_, (float)params.sor_omega, b_.p, pin, pout);
                        else
                            k_rbsor4<2, kStPlain><<<tls_.nb, kThreads, 0, st_>>>(
                                g_, tls_, (float)params.sor_omega, b_.p, pin, pout);
                        LAUNCH_CHECK();
                    }
                    if (sweeps & 1)
                        HIPT(hipMemcpyAsync(p_.p, p2_.p, n_ * sizeof(T), hipMemcpyDeviceToDevice,
                                            st_));
                    last_kind_ = AETHER_SOLVER_RBGS;
                    last_sweeps_ = sweeps;
                    return;
                }
                for (int it = 0; it < sweeps; ++it)
                    for (int color = 0; color < 2; ++color) {
                        if (wt_ok())
                            VZU_DISPATCH(vzu_, (k_rbgs4<VZU, kStWT><<<tlv_.nb, kThreads, 0, st_>>>(
                                                   g_, tlv_, color, (float)params.sor_omega, b_.p,
                                                   p_.p)));
                        else
                            VZU_DISPATCH(vzu_, (k_rbgs4<VZU, kStPlain><<<tlv_.nb, kThreads, 0, st_>>>(
                                                   g_, tlv_, color, (float)params.sor_omega, b_.p,
                                                   p_.p)));
                        LAUNCH_CHECK();
                    }
            }
            last_kind_ = AETHER_SOLVER_RBGS;
            last_sweeps_ = sweeps;
            return;
        }
        if constexpr (sizeof(T) == 8) {
            // FP64 colour passes on the two-wide layout (k_rbgs2: k_rbgs's
            // per-cell arithmetic, bit-identical)
            if (tlv_.nb > 0 && params.cg_form != AETHER_CG_SCALAR) {
                const bool wt = n_ * sizeof(T) < (size_t(1) << 31);
                if (tlv_.gx == 1 && 64 % tlv_.txv == 0 && sweeps > 0) {
                    // one launch per sweep (k_rbsor2), 2 planes per thread,
                    // from p to a second array and back
                    VTiling ts = tlv_;
                    ts.gz = (g_.nz + 1) / 2;
                    ts.nb = ts.gx * ts.gy * ts.gz;
                    ts.remap = (ts.nb % 8 == 0) ? 1 : 0;
                    if (!p2_.p) p2_.alloc(n_);
                    for (int it = 0; it < sweeps; ++it) {
                        const T *pin = (it & 1) ? p2_.p : p_.p;
                        T *pout = (it & 1) ? p_.p : p2_.p;
                        if (wt)
                            k_rbsor2<2, kStWT><<<ts.nb, kThreads, 0, st_>>>(
                                g_, ts, params.sor_omega, b_.p, pin, pout);
                        else
                            k_rbsor2<2, kStPlain><<<ts.nb, kThreads, 0, st_>>>(
                                g_, ts, params.sor_omega, b_.p, pin, pout);
                        LAUNCH_CHECK();
                    }
                    if (sweeps & 1)
                        HIPT(hipMemcpyAsync(p_.p, p2_.p, n_ * sizeof(T), hipMemcpyDeviceToDevice,
                                            st_));
                    last_kind_ = AETHER_SOLVER_RBGS;
                    last_sweeps_ = sweeps;
                    return;
                }
                for (int it = 0; it < sweeps; ++it)
                    for (int color = 0; color < 2; ++color) {
                        if (wt)
                            VZU_DISPATCH(vzu_, (k_rbgs2<VZU, kStWT><<<tlv_.nb, kThreads, 0, st_>>>(
                                                   g_, tlv_, color, params.sor_omega, b_.p, p_.p)));
                        else
                            VZU_DISPATCH(vzu_, (k_rbgs2<VZU, kStPlain><<<tlv_.nb, kThreads, 0,
                                                                         st_>>>(
                                                   g_, tlv_, color, params.sor_omega, b_.p, p_.p)));
                        LAUNCH_CHECK();
                    }
                last_kind_ = AETHER_SOLVER_RBGS;
                last_sweeps_ = sweeps;
                return;
            }
        }
        for (int it = 0; it < sweeps; ++it)
            for (int color = 0; color < 2; ++color) {
                k_rbgs<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, color, (T)params.sor_omega,
                                                             b_.p, p_.p);
                LAUNCH_CHECK();
            }
        last_kind_ = AETHER_SOLVER_RBGS;
        last_sweeps_ = sweeps;
    }

    // Mixed-precision projection: b and p in FP64, the FP32 CG solves the
    // correction equation A e = r to ir_inner_rel_tol, p += e, until the FP64
    // residual gives max|div u| = max|b - A p| / h^2 <= ir_div_tol.  One
    // host synchronisation per refinement sweep reads max|r|.
    void project_ir(int step_index) {
        k_div64<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, h_, cur_[0], cur_[1], cur_[2], b64_.p);
        LAUNCH_CHECK();
        const double inv_h2 = 1.0 / (h_ * h_);
        int outer = 0, inner = 0;
        double rmax = 0.0;
        for (;; ++outer) {
            k_residual64<T><<<red_grid_, cell_blk_, 0, st_>>>(g_, zchunk_, b64_.p, p64_.p, b_.p,
                                                              partA_.p, partD64_.p);
            LAUNCH_CHECK();
            k_partials_final<<<1, kThreads, 0, st_>>>(partA_.p, nb_, 1, scalar_.p);
            LAUNCH_CHECK();
            rmax = read_scalar();
            if (rmax * inv_h2 <= params.ir_div_tol || outer >= params.ir_max_outer) break;
            zero(p_.p);  // inner unknown e starts at 0
            rel_override_ = params.ir_inner_rel_tol;
            ZU_DISPATCH(zu_, (k_div_cginit<T, ZU><<<tl_.nb, cell_blk_, 0, st_>>>(
                                  g_, tl_, (T)h_, 0, cur_[0], cur_[1], cur_[2], p_.p, b_.p, r_.p,
                                  partB_.p, partC_.p, cg_.p)));
            LAUNCH_CHECK();
            enqueue_cg_loop(params.solver_iters, -1);
            rel_override_ = -1.0;
            CgState cs;
            HIPT(hipMemcpyAsync(&cs, cg_.p, sizeof(cs), hipMemcpyDeviceToHost, st_));
            k_add64<T><<<blocks_for(n_), kThreads, 0, st_>>>(p_.p, (long)n_, p64_.p);
            LAUNCH_CHECK();
            sync();
            inner += cs.iters;
        }
        k_gradient64_bc<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, 1.0 / h_, p64_.p, cur_[0],
                                                               cur_[1], cur_[2]);
        LAUNCH_CHECK();
        ir_outer_ = outer;
        ir_inner_ = inner;
        ir_div_ = rmax * inv_h2;
        if (step_index >= 0) ir_log_.push_back(inner);
        last_kind_ = AETHER_SOLVER_CG;
    }

    // defer_gradient: leave the closing gradient subtraction to the scalar
    // advection pass that follows (enqueue_advect_scalars with_gradient);
    // returns whether it was left (never for the mixed-precision projection)
    bool enqueue_project(int step_index, bool defer_gradient = false) {
        sync_tilings();
        if (ir64_) {
            project_ir(step_index);
            return false;
        }
        // src/fluid.rs:496-529; the opening wall BC is applied through masked
        // reads in the divergence, the closing one in the gradient kernel
        if (params.solver_kind == AETHER_SOLVER_CG) {
            ZU_DISPATCH(zu_, (k_div_cginit<T, ZU><<<tl_.nb, cell_blk_, 0, st_>>>(
                                  g_, tl_, (T)h_, 1, cur_[0], cur_[1], cur_[2], p_.p, b_.p,
                                  r_.p, partB_.p, partC_.p, cg_.p)));
            LAUNCH_CHECK();
            enqueue_cg_loop(params.solver_iters, step_index);
        } else {
            k_divergence<T><<<red_grid_, cell_blk_, 0, st_>>>(g_, zchunk_, (T)h_, (T)(1.0 / h_),
                                                               1, cur_[0], cur_[1], cur_[2],
                                                               b_.p, partA_.p);
            LAUNCH_CHECK();
            enqueue_rbgs(params.solver_iters);
        }
        if (defer_gradient) return true;
        k_gradient_bc<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, (T)(1.0 / h_), p_.p, cur_[0],
                                                             cur_[1], cur_[2]);
        LAUNCH_CHECK();
        return false;
    }

    void project() override {
        enqueue_project(-1);
        sync();
    }

    void solve_cg(int iters) override {
        if (iters < 0) raise(AETHER_EINVAL, "iterations must be >= 0");
        sync_tilings();
        ZU_DISPATCH(zu_, (k_div_cginit<T, ZU><<<tl_.nb, cell_blk_, 0, st_>>>(
                              g_, tl_, (T)h_, 0, cur_[0], cur_[1], cur_[2], p_.p, b_.p, r_.p,
                              partB_.p, partC_.p, cg_.p)));
        LAUNCH_CHECK();
        enqueue_cg_loop(iters, -1);
        sync();
    }

    void solve_rbgs(int sweeps) override {
        if (sweeps < 0) raise(AETHER_EINVAL, "sweeps must be >= 0");
        enqueue_rbgs(sweeps);
        sync();
    }

    void enqueue_sources(double t, double dt) {
        if (patch_on_) {  // rollout inflow: density and temperature set from the controls
            const int cnt = patch_n_.x * patch_n_.y * patch_n_.z;
            k_patch_set<T><<<(cnt + kThreads - 1) / kThreads, cell_blk_, 0, st_>>>(
                g_, patch_lo_, patch_n_, ctrl_.p, cur_[3], cur_[4]);
            LAUNCH_CHECK();
            return;
        }
        if (params.source_kind == AETHER_SOURCE_JETS) {
            // scalar set-up of src/fluid.rs:533-563 on the host (libm sin/cos)
            const double TAU = 6.28318530717958647692528676655900577;
            const double cx = 0.5 * (double)g_.nx, cz = 0.5 * (double)g_.nz;
            const int base_j = std::min(3, g_.ny - 3);
            const double pulse = 0.65 + 0.35 * std::sin(t * TAU * 0.55) +
                                 0.25 * std::fmax(std::cos(t * TAU * 1.35), 0.0);
            const double angle = t * 1.85;
            const double orbit_r = (double)g_.nx * 0.14;
            JetParams jp;
            const double ca = std::cos(angle), sa = std::sin(angle);
            const double j0[6] = {cx + orbit_r * ca, (double)base_j + 1.5, cz + orbit_r * sa,
                                  -2.6 * sa, 7.2 * pulse, 2.6 * ca};
            const double j1[6] = {cx - orbit_r * ca, (double)base_j + 2.0, cz - orbit_r * sa,
                                  2.6 * sa, 6.4 * pulse, -2.6 * ca};
            std::memcpy(jp.jet[0], j0, sizeof(j0));
            std::memcpy(jp.jet[1], j1, sizeof(j1));
            const double r_cells = std::fmax((double)g_.nx * 0.105, 2.2);
            jp.r2 = r_cells * r_cells;
            jp.pulse = pulse;
            jp.dt = dt;
            int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
            const int nn[3] = {g_.nx, g_.ny, g_.nz};
            for (int q = 0; q < 2; ++q) {
                int qlo[3], qhi[3];
                bool empty = false;
                for (int d = 0; d < 3; ++d) {
                    const double c = jp.jet[q][d];
                    qlo[d] = (int)std::fmax(std::floor(c - r_cells - 1.0), 1.0);
                    qhi[d] = (int)std::fmin(std::ceil(c + r_cells + 1.0), (double)(nn[d] - 2));
                    if (qlo[d] > qhi[d]) empty = true;
                }
                if (empty) continue;
                for (int d = 0; d < 3; ++d) {
                    lo[d] = std::min(lo[d], qlo[d]);
                    hi[d] = std::max(hi[d], qhi[d]);
                }
            }
            if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return;
            const int ni = hi[0] - lo[0] + 1, nj = hi[1] - lo[1] + 1, nk = hi[2] - lo[2] + 1;
            const unsigned nbk = (unsigned)(((long)ni * nj * nk + kThreads - 1) / kThreads);
            k_jets<T><<<nbk, cell_blk_, 0, st_>>>(g_, jp, lo[0], lo[1], lo[2], ni, nj, nk,
                                                  cur_[0], cur_[1], cur_[2], cur_[3], cur_[4]);
            LAUNCH_CHECK();
        } else if (params.source_kind == AETHER_SOURCE_SPHERE) {
            const double r = params.source_radius;
            if (!(r > 0.0)) return;
            const int nn[3] = {g_.nx, g_.ny, g_.nz};
            int lo[3], hi[3];
            for (int d = 0; d < 3; ++d) {
                const double c = params.source_center[d];
                lo[d] = std::max(1, (int)std::floor(c - r));
                hi[d] = std::min(nn[d] - 2, (int)std::ceil(c + r));
                if (hi[d] < lo[d]) return;
            }
            const int ni = hi[0] - lo[0] + 1, nj = hi[1] - lo[1] + 1, nk = hi[2] - lo[2] + 1;
            const unsigned nbk = (unsigned)(((long)ni * nj * nk + kThreads - 1) / kThreads);
            k_sphere<T><<<nbk, cell_blk_, 0, st_>>>(
                g_, params.source_center[0], params.source_center[1], params.source_center[2],
                r * r, params.source_density, params.source_temperature, lo[0], lo[1], lo[2], ni,
                nj, nk, cur_[3], cur_[4]);
            LAUNCH_CHECK();
        }
    }

    void inject_sources(double t, double dt) override {
        enqueue_sources(t, dt);
        sync();
    }

    void enqueue_telemetry(double t, double *dst_dev) {
        k_telemetry_part<T><<<red_grid_, cell_blk_, 0, st_>>>(g_, zchunk_, h_, cur_[0], cur_[1],
                                                              cur_[2], cur_[3], part_tel_.p,
                                                              nonfinite_.p);
        LAUNCH_CHECK();
        k_telemetry_final<<<1, cell_blk_, 0, st_>>>(part_tel_.p, nb_, t, dst_dev);
        LAUNCH_CHECK();
    }

    static void unpack(const double *raw, aether_fluid_telemetry *out) {
        out->time_sec = raw[0];
        out->kinetic_energy = raw[1];
        out->enstrophy = raw[2];
        out->peak_vorticity = raw[3];
        out->total_mass = raw[4];
        out->center_of_mass[0] = raw[5];
        out->center_of_mass[1] = raw[6];
        out->center_of_mass[2] = raw[7];
    }

    void telemetry(double t, aether_fluid_telemetry *out) override {
        enqueue_telemetry(t, tel_.p);
        double raw[8];
        HIPT(hipMemcpyAsync(raw, tel_.p, sizeof(raw), hipMemcpyDeviceToHost, st_));
        sync();
        if (out) unpack(raw, out);
    }

    // src/fluid.rs:598-611.  The velocity advections share passes, a sphere
    // source rides in the curl pass, and the projection's gradient in the
    // scalar advection's first pass.
    void enqueue_step(double dt, double t, double *tel_dst, int step_index) {
        enqueue_advect_velocity(dt, params.dissipation_velocity);
        const bool sphere = !patch_on_ && params.source_kind == AETHER_SOURCE_SPHERE;
        if (!sphere) enqueue_sources(t, dt);
        enqueue_buoyancy_confinement(dt, sphere);
        const bool deferred = enqueue_project(step_index, true);
        enqueue_advect_scalars(dt, deferred);
        enqueue_telemetry(t, tel_dst);
    }

    void run(double dt, const double *times, int nsteps, aether_fluid_telemetry *out) override {
        if (nsteps < 0) raise(AETHER_EINVAL, "nsteps must be >= 0");
        if (nsteps == 0) return;
        if (tel_ring_.n < (size_t)nsteps * 8) tel_ring_.alloc((size_t)nsteps * 8);
        if (iters_dev_.n < (size_t)nsteps) iters_dev_.alloc((size_t)nsteps);
        if (timer.enabled) timer.reset();
        ir_log_.clear();
        for (int s = 0; s < nsteps; ++s)
            enqueue_step(dt, times[s], tel_ring_.p + (size_t)s * 8, s);
        std::vector<double> raw((size_t)nsteps * 8);
        HIPT(hipMemcpyAsync(raw.data(), tel_ring_.p, raw.size() * sizeof(double),
                            hipMemcpyDeviceToHost, st_));
        std::vector<int> iters(nsteps);
        HIPT(hipMemcpyAsync(iters.data(), iters_dev_.p, nsteps * sizeof(int),
                            hipMemcpyDeviceToHost, st_));
        sync();
        if (out)
            for (int s = 0; s < nsteps; ++s) unpack(raw.data() + (size_t)s * 8, out + s);
        if (ir64_) iters.assign(ir_log_.begin(), ir_log_.end());
        if (params.solver_kind == AETHER_SOLVER_RBGS)
            std::fill(iters.begin(), iters.end(), params.solver_iters);
        last_run_iters_ = iters;
        if (timer.enabled) {
            // count only launches that did work: iterations < iters of that step
            for (const auto &r : timer.recs) {
                if (r.which != 2 && r.it >= iters[r.step]) continue;
                float ms = 0;
                HIPT(hipEventElapsedTime(&ms, r.a, r.b));
                timer.sum_ms[r.which] += ms;
                timer.count[r.which] += 1;
            }
        }
    }

    T *field_ptr(int field) {
        switch (field) {
        case AETHER_FIELD_U:
        case AETHER_FIELD_V:
        case AETHER_FIELD_W:
        case AETHER_FIELD_DENSITY:
        case AETHER_FIELD_TEMPERATURE:
            return cur_[field];
        case AETHER_FIELD_PRESSURE:
            return p_.p;
        case AETHER_FIELD_DIVERGENCE:
            return b_.p;
        default:
            raise(AETHER_EINVAL, "unknown field id " + std::to_string(field));
        }
    }

    double *field64(int field) {
        if (!ir64_) return nullptr;
        if (field == AETHER_FIELD_PRESSURE) return p64_.p;
        if (field == AETHER_FIELD_DIVERGENCE) return b64_.p;
        return nullptr;
    }

    void upload(int field, const double *src) override {
        if (double *d64 = field64(field)) {
            HIPT(hipMemcpyAsync(d64, src, n_ * sizeof(double), hipMemcpyHostToDevice, st_));
            sync();
            return;
        }
        upload_to(src, field_ptr(field));
        sync();
    }

    // n_ f64 host values into a device array of T (converted on the device);
    // enqueued, with a later synchronisation.
    void upload_to(const double *src, T *dst) {
        if constexpr (sizeof(T) == 8) {
            HIPT(hipMemcpyAsync(dst, src, n_ * sizeof(double), hipMemcpyHostToDevice, st_));
        } else {
            if (!staging_.p) staging_.alloc(n_);
            HIPT(hipMemcpyAsync(staging_.p, src, n_ * sizeof(double), hipMemcpyHostToDevice,
                                st_));
            k_convert<double, T><<<blocks_for(n_), kThreads, 0, st_>>>(staging_.p, (long)n_,
                                                                         dst);
            LAUNCH_CHECK();
        }
        // the staging buffer is reused by the next call: wait for this copy
        if constexpr (sizeof(T) != 8) sync();
    }

    // A device array of T into n_ f64 host values (converted on the device).
    void download_to(const T *src, double *dst) {
        if constexpr (sizeof(T) == 8) {
            HIPT(hipMemcpyAsync(dst, src, n_ * sizeof(double), hipMemcpyDeviceToHost, st_));
        } else {
            if (!staging_.p) staging_.alloc(n_);
            k_convert<T, double><<<blocks_for(n_), kThreads, 0, st_>>>(src, (long)n_,
                                                                         staging_.p);
            LAUNCH_CHECK();
            HIPT(hipMemcpyAsync(dst, staging_.p, n_ * sizeof(double), hipMemcpyDeviceToHost,
                                st_));
        }
        sync();
    }

    void download(int field, double *dst) override {
        if (double *d64 = field64(field)) {
            HIPT(hipMemcpyAsync(dst, d64, n_ * sizeof(double), hipMemcpyDeviceToHost, st_));
            sync();
            return;
        }
        download_to(field_ptr(field), dst);
    }

    // src/fluid.rs:94-128 on a caller's f64 field in this grid's layout,
    // sampled in FP64 whatever the solver's precision.
    void sample_field(const double *field, const double *xyz, size_t npts,
                      double *out) override {
        if (npts == 0) return;
        DevBuf<double> f(n_), dxyz(npts * 3), dout(npts);
        HIPT(hipMemcpyAsync(f.p, field, n_ * sizeof(double), hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(dxyz.p, xyz, npts * 3 * sizeof(double), hipMemcpyHostToDevice, st_));
        k_sample_points<double><<<(unsigned)((npts + kThreads - 1) / kThreads), cell_blk_, 0,
                                  st_>>>(g_, f.p, dxyz.p, (long)npts, dout.p);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(out, dout.p, npts * sizeof(double), hipMemcpyDeviceToHost, st_));
        sync();
    }

    void sample(int field, const double *xyz, size_t npts, double *out) override {
        T *f = field_ptr(field);
        if (npts == 0) return;
        DevBuf<double> dxyz(npts * 3), dout(npts);
        HIPT(hipMemcpyAsync(dxyz.p, xyz, npts * 3 * sizeof(double), hipMemcpyHostToDevice, st_));
        k_sample_points<T><<<(unsigned)((npts + kThreads - 1) / kThreads), cell_blk_, 0, st_>>>(
            g_, f, dxyz.p, (long)npts, dout.p);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(out, dout.p, npts * sizeof(double), hipMemcpyDeviceToHost, st_));
        sync();
    }

    std::unique_ptr<MediumBase> make_medium(const double *lo, const double *hi) override {
        auto m = std::make_unique<Medium<T>>(g_, cur_[3], cur_[4], lo, hi, st_);
        m->max_density = max_density();
        return m;
    }

    double max_density() override {
        // src/fluid.rs:704-707: max(0, all density entries)
        const unsigned nbk = partial_blocks(n_, partA_);
        k_max_part<T><<<nbk, cell_blk_, 0, st_>>>(cur_[3], (long)n_, partA_.p);
        LAUNCH_CHECK();
        k_partials_final<<<1, kThreads, 0, st_>>>(partA_.p, (int)nbk, 1, scalar_.p);
        LAUNCH_CHECK();
        return read_scalar();
    }

    void stats(aether_fluid_solve_stats *out) override {
        std::memset(out, 0, sizeof(*out));
        if (last_kind_ == AETHER_SOLVER_RBGS) {
            out->iterations = last_sweeps_;
            out->converged = 1;
            return;
        }
        CgState cs;
        HIPT(hipMemcpyAsync(&cs, cg_.p, sizeof(cs), hipMemcpyDeviceToHost, st_));
        sync();
        out->iterations = cs.iters;
        out->converged = cs.converged;
        out->b_norm2 = cs.b2;
        out->r_norm2_initial = cs.r2_init;
        out->r_norm2_final = cs.r2_final;
        if (ir64_) {
            out->iterations = ir_inner_;
            out->outer_iterations = ir_outer_;
            out->div_max_final = ir_div_;
            out->converged = ir_div_ <= params.ir_div_tol;
        }
    }


    // ---------------------------------------------------------------------
    // Differentiable rollouts (checkpointed forward + hand-written adjoint)
    // ---------------------------------------------------------------------
    struct StepRec {
        DevBuf<T> u0, v0, w0, u1, v1, w1, rs, ts, u3, v3, w3;
    };
    struct Ckpt {
        DevBuf<T> f[5], p;
    };

    void copy_arr(const T *a, T *b) {
        k_copy<T><<<blocks_for(n_), kThreads, 0, st_>>>(a, (long)n_, b);
        LAUNCH_CHECK();
    }

    void save_ckpt(Ckpt &c) {
        for (int a = 0; a < 5; ++a) {
            if (!c.f[a].p) c.f[a].alloc(n_);
            copy_arr(cur_[a], c.f[a].p);
        }
        if (!c.p.p) c.p.alloc(n_);
        copy_arr(p_.p, c.p.p);
    }

    void load_ckpt(const Ckpt &c) {
        for (int a = 0; a < 5; ++a) copy_arr(c.f[a].p, cur_[a]);
        copy_arr(c.p.p, p_.p);
    }

    void step_plain(double dt) {
        enqueue_advect1(0, dt, params.dissipation_velocity);
        enqueue_advect1(1, dt, params.dissipation_velocity);
        enqueue_advect1(2, dt, params.dissipation_velocity);
        enqueue_sources(0.0, dt);
        enqueue_buoyancy_confinement(dt);
        enqueue_project(-1);
        enqueue_advect_scalars(dt);
    }

    // The same step as step_plain, keeping what the adjoint step needs.
    void step_recorded(double dt, StepRec &r) {
        for (DevBuf<T> *b : {&r.u0, &r.v0, &r.w0, &r.u1, &r.v1, &r.w1, &r.rs, &r.ts, &r.u3,
                             &r.v3, &r.w3})
            if (!b->p) b->alloc(n_);
        copy_arr(cur_[0], r.u0.p);
        copy_arr(cur_[1], r.v0.p);
        copy_arr(cur_[2], r.w0.p);
        enqueue_advect1(0, dt, params.dissipation_velocity);
        enqueue_advect1(1, dt, params.dissipation_velocity);
        enqueue_advect1(2, dt, params.dissipation_velocity);
        copy_arr(cur_[0], r.u1.p);
        copy_arr(cur_[1], r.v1.p);
        copy_arr(cur_[2], r.w1.p);
        enqueue_sources(0.0, dt);
        copy_arr(cur_[3], r.rs.p);
        copy_arr(cur_[4], r.ts.p);
        enqueue_buoyancy_confinement(dt);
        enqueue_project(-1);
        copy_arr(cur_[0], r.u3.p);
        copy_arr(cur_[1], r.v3.p);
        copy_arr(cur_[2], r.w3.p);
        enqueue_advect_scalars(dt);
    }

    // Adjoint of one MacCormack advection: lam_out -> lam_src, lam_U/V/W
    // (added to).  The scatters accumulate in 64-bit fixed point (order-
    // independent, so the gradient is reproducible bit for bit), scaled from
    // |lam_out|_1: every contribution of the correction and the forward pass
    // together sums in absolute value to at most 2.5 |diss| |lam_out|_1
    // (the correction passes |diss lam| to the forward adjoint plus
    // 0.5 |diss lam| through the trilinear weights, which sum to 1, and at
    // most |diss lam| to the source; the forward pass passes its adjoint on
    // through weights summing to 1).  lam_src may be the same array as one
    // of lam_U/V/W: the flush adds in a fixed order per cell.
    void adv_adjoint(double dt, const T *src, const T *U, const T *V, const T *W, double diss,
                     int clamp0, const T *lam_out, T *lam_src, T *lam_U, T *lam_V, T *lam_W) {
        const T s = (T)(dt / h_);
        k_sl_forward<T, 1><<<cell_grid_, cell_blk_, 0, st_>>>(g_, s, U, V, W, src, nullptr,
                                                              fwd_[0].p, nullptr);
        LAUNCH_CHECK();
        const unsigned nbk = partial_blocks(n_, adj_part_);
        k_abs_sum<T><<<nbk, kThreads, 0, st_>>>(lam_out, (long)n_, adj_part_.p);
        LAUNCH_CHECK();
        k_fix_scale<<<1, kThreads, 0, st_>>>(adj_part_.p, (int)nbk, 2.5 * std::fabs(diss),
                                             fix_.p);
        LAUNCH_CHECK();
        k_fill<fixacc_t><<<blocks_for(n_), kThreads, 0, st_>>>(acc_src_.p, (long)n_, 0);
        LAUNCH_CHECK();
        k_fill<fixacc_t><<<blocks_for(n_), kThreads, 0, st_>>>(acc_fwd_.p, (long)n_, 0);
        LAUNCH_CHECK();
        const dim3 grid((g_.nx + WTX - 1) / WTX, (g_.ny + WTY - 1) / WTY,
                        (g_.nz + WTZ - 1) / WTZ);
        // LDS-windowed scatters (adjoint_kernels.hpp) once the grid has a tile
        // per CU (256); smaller grids keep the one-cell-per-thread launches,
        // which spread over more workgroups (measured with float atomics:
        // 128^3 fwd+adj/fwd 3.23 -> 2.27 windowed; 64^3 2.74 direct, 2.83
        // windowed)
        if ((long)grid.x * grid.y * grid.z >= 256) {
            const dim3 blk(WTX, WTY);
            k_adj_mac_correct_w<T><<<grid, blk, 0, st_>>>(
                g_, s, (T)diss, clamp0, src, fwd_[0].p, U, V, W, lam_out, fix_.p, acc_src_.p,
                acc_fwd_.p, adj_tu_[0].p, adj_tu_[1].p, adj_tu_[2].p);
            LAUNCH_CHECK();
            k_adj_sl_forward_w<T><<<grid, blk, 0, st_>>>(g_, s, src, U, V, W, fix_.p, acc_fwd_.p,
                                                         acc_src_.p, adj_tu_[0].p, adj_tu_[1].p,
                                                         adj_tu_[2].p);
            LAUNCH_CHECK();
        } else {
            k_adj_mac_correct<T><<<cell_grid_, cell_blk_, 0, st_>>>(
                g_, s, (T)diss, clamp0, src, fwd_[0].p, U, V, W, lam_out, fix_.p, acc_src_.p,
                acc_fwd_.p, adj_tu_[0].p, adj_tu_[1].p, adj_tu_[2].p);
            LAUNCH_CHECK();
            k_adj_sl_forward<T><<<cell_grid_, cell_blk_, 0, st_>>>(
                g_, s, src, U, V, W, fix_.p, acc_fwd_.p, acc_src_.p, adj_tu_[0].p, adj_tu_[1].p,
                adj_tu_[2].p);
            LAUNCH_CHECK();
        }
        k_adj_flush<T><<<blocks_for(n_), kThreads, 0, st_>>>((long)n_, fix_.p, acc_src_.p,
                                                             adj_tu_[0].p, adj_tu_[1].p,
                                                             adj_tu_[2].p, lam_src, lam_U, lam_V,
                                                             lam_W);
        LAUNCH_CHECK();
    }

    // P^T = P: the forward projection applied to (lu, lv, lw) with its own
    // pressure buffer (warm-started by the previous adjoint step).
    void project_arrays(T *lu, T *lv, T *lw, int iters, double rel_tol) {
        T *save[3] = {cur_[0], cur_[1], cur_[2]};
        cur_[0] = lu;
        cur_[1] = lv;
        cur_[2] = lw;
        std::swap(p_.p, adj_p_.p);
        const int si = params.solver_iters;
        const double st = params.cg_rel_tol;
        const int sk = params.solver_kind;
        params.solver_iters = iters;
        params.cg_rel_tol = rel_tol;
        params.solver_kind = AETHER_SOLVER_CG;
        enqueue_project(-1);
        params.solver_iters = si;
        params.cg_rel_tol = st;
        params.solver_kind = sk;
        std::swap(p_.p, adj_p_.p);
        for (int a = 0; a < 3; ++a) cur_[a] = save[a];
    }

    void adjoint_step(double dt, StepRec &r, const aether_rollout_desc &d, int step_slot) {
        // a) density / temperature advection (velocity u3)
        zero(adj_b_[0].p);
        zero(adj_b_[1].p);
        adv_adjoint(dt, r.rs.p, r.u3.p, r.v3.p, r.w3.p, params.dissipation_density, 1,
                    adj_ar_.p, adj_b_[0].p, adj_a_[0].p, adj_a_[1].p, adj_a_[2].p);
        adv_adjoint(dt, r.ts.p, r.u3.p, r.v3.p, r.w3.p, params.dissipation_temperature, 1,
                    adj_at_.p, adj_b_[1].p, adj_a_[0].p, adj_a_[1].p, adj_a_[2].p);
        // b) projection (self-adjoint)
        project_arrays(adj_a_[0].p, adj_a_[1].p, adj_a_[2].p, d.adjoint_cg_iters,
                       d.adjoint_cg_rel_tol);
        // c) confinement + buoyancy + wall masks around velocity u1
        const double inv_2h = 0.5 / h_;
        const SphereSrc none = {0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        k_curl_mag<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, (T)inv_2h, r.u1.p, r.v1.p, r.w1.p,
                                                          curl_.p, none, nullptr, nullptr);
        LAUNCH_CHECK();
        for (int a = 0; a < 3; ++a) zero(adj_c_[a].p);
        k_adj_confine1<T><<<cell_grid_, cell_blk_, 0, st_>>>(
            g_, (T)dt, (T)inv_2h, (T)(params.vorticity_eps * h_), (T)params.buoyancy_beta,
            (T)params.gravity_alpha, r.u1.p, r.v1.p, r.w1.p, curl_.p, r.ts.p, adj_a_[0].p,
            adj_a_[1].p, adj_a_[2].p, adj_c_[0].p, adj_c_[1].p, adj_c_[2].p, adj_lw_[0].p,
            adj_lw_[1].p, adj_lw_[2].p, adj_lg_[0].p, adj_lg_[1].p, adj_lg_[2].p, adj_b_[0].p,
            adj_b_[1].p, adj_part_.p);
        LAUNCH_CHECK();
        const int ncell = (int)(cell_grid_.x * cell_grid_.y * cell_grid_.z);
        // this step's d loss / d beta, kept on the device (summed over the
        // steps in reverse order at the end: no host synchronisation per step)
        k_partials_final<<<1, kThreads, 0, st_>>>(adj_part_.p, ncell, 0, gbeta_.p + step_slot);
        LAUNCH_CHECK();
        k_adj_confine2<T><<<cell_grid_, cell_blk_, 0, st_>>>(
            g_, (T)inv_2h, r.u1.p, r.v1.p, r.w1.p, curl_.p, adj_lg_[0].p, adj_lg_[1].p,
            adj_lg_[2].p, adj_lw_[0].p, adj_lw_[1].p, adj_lw_[2].p);
        LAUNCH_CHECK();
        k_adj_confine3<T><<<cell_grid_, cell_blk_, 0, st_>>>(g_, (T)inv_2h, adj_lw_[0].p,
                                                            adj_lw_[1].p, adj_lw_[2].p,
                                                            adj_c_[0].p, adj_c_[1].p, adj_c_[2].p);
        LAUNCH_CHECK();
        // d) inflow patch
        const int cnt = patch_n_.x * patch_n_.y * patch_n_.z;
        k_patch_adj<T><<<(cnt + kThreads - 1) / kThreads, cell_blk_, 0, st_>>>(
            g_, patch_lo_, patch_n_, adj_b_[0].p, adj_b_[1].p, grad_c_.p);
        LAUNCH_CHECK();
        // e-g) velocity advections, last to first: w (vel u1, v1, w0),
        // v (vel u1, v0, w0), u (vel u0, v0, w0)
        for (int a = 0; a < 3; ++a) zero(adj_d_[a].p);
        const double dv = params.dissipation_velocity;
        adv_adjoint(dt, r.w0.p, r.u1.p, r.v1.p, r.w0.p, dv, 0, adj_c_[2].p, adj_d_[2].p,
                    adj_c_[0].p, adj_c_[1].p, adj_d_[2].p);
        adv_adjoint(dt, r.v0.p, r.u1.p, r.v0.p, r.w0.p, dv, 0, adj_c_[1].p, adj_d_[1].p,
                    adj_c_[0].p, adj_d_[1].p, adj_d_[2].p);
        adv_adjoint(dt, r.u0.p, r.u0.p, r.v0.p, r.w0.p, dv, 0, adj_c_[0].p, adj_d_[0].p,
                    adj_d_[0].p, adj_d_[1].p, adj_d_[2].p);
        for (int a = 0; a < 3; ++a) std::swap(adj_a_[a].p, adj_d_[a].p);
        std::swap(adj_ar_.p, adj_b_[0].p);
        std::swap(adj_at_.p, adj_b_[1].p);
    }

    // Checks a rollout description and loads its controls; sets the patch
    // source and beta.  Returns the number of control cells.
    size_t rollout_setup(const aether_rollout_desc &d, const double *ctrl, double beta) {
        if (ir64_) raise(AETHER_EINVAL, "rollouts need AETHER_FP32 or AETHER_FP64 precision");
        if (d.nsteps < 1 || d.checkpoint_every < 1) raise(AETHER_EINVAL, "bad rollout lengths");
        const int nn[3] = {g_.nx, g_.ny, g_.nz};
        for (int a = 0; a < 3; ++a)
            if (d.patch_lo[a] < 1 || d.patch_hi[a] > nn[a] - 1 || d.patch_hi[a] <= d.patch_lo[a])
                raise(AETHER_EINVAL, "patch must be a non-empty box of interior cells");
        patch_lo_ = make_int3(d.patch_lo[0], d.patch_lo[1], d.patch_lo[2]);
        patch_n_ = make_int3(d.patch_hi[0] - d.patch_lo[0], d.patch_hi[1] - d.patch_lo[1],
                             d.patch_hi[2] - d.patch_lo[2]);
        const size_t cnt = (size_t)patch_n_.x * patch_n_.y * patch_n_.z;
        if (ctrl_.n < cnt) ctrl_.alloc(cnt);
        if (grad_c_.n < cnt) grad_c_.alloc(cnt);
        HIPT(hipMemcpyAsync(ctrl_.p, ctrl, cnt * sizeof(double), hipMemcpyHostToDevice, st_));
        params.buoyancy_beta = beta;
        patch_on_ = true;
        return cnt;
    }

    void ensure_adjoint_buffers() {
        for (DevBuf<T> *b : {&adj_a_[0], &adj_a_[1], &adj_a_[2], &adj_ar_, &adj_at_, &adj_b_[0],
                             &adj_b_[1], &adj_c_[0], &adj_c_[1], &adj_c_[2], &adj_d_[0],
                             &adj_d_[1], &adj_d_[2], &adj_tu_[0], &adj_tu_[1], &adj_tu_[2],
                             &adj_lw_[0], &adj_lw_[1], &adj_lw_[2], &adj_lg_[0], &adj_lg_[1],
                             &adj_lg_[2], &adj_p_})
            if (!b->p) b->alloc(n_);
        for (DevBuf<fixacc_t> *b : {&acc_src_, &acc_fwd_})
            if (!b->p) b->alloc(n_);
        if (!fix_.p) fix_.alloc(1);
        if (!adj_part_.p) adj_part_.alloc(cell_grid_.x * cell_grid_.y * cell_grid_.z);
    }

    // Forward sweep of d.nsteps steps from the current state, saving a
    // checkpoint every d.checkpoint_every steps; with `record`, the last
    // segment's steps are recorded for the reverse sweep.  Leaves the final
    // state in the solver.
    void rollout_forward_sweep(const aether_rollout_desc &d, bool record) {
        const int K = d.checkpoint_every;
        const int nseg = (d.nsteps + K - 1) / K;
        while ((int)ckpts_.size() < nseg) ckpts_.push_back(std::make_unique<Ckpt>());
        const int last0 = (nseg - 1) * K;
        if (record)
            while ((int)recs_.size() < K) recs_.push_back(std::make_unique<StepRec>());
        for (int t = 0; t < d.nsteps; ++t) {
            if (t % K == 0) save_ckpt(*ckpts_[t / K]);
            if (record && t >= last0)
                step_recorded(d.dt, *recs_[t - last0]);
            else
                step_plain(d.dt);
        }
    }

    // Reverse sweep: from the adjoints of the final state in adj_a_ (u, v,
    // w), adj_ar_ (density) and adj_at_ (temperature) to those of the
    // initial state (same arrays), the controls (grad_c_) and beta.  Needs
    // the checkpoints and the recorded last segment of the forward sweep.
    double rollout_reverse_sweep(const aether_rollout_desc &d, size_t cnt) {
        const int K = d.checkpoint_every;
        const int nseg = (d.nsteps + K - 1) / K;
        zero(adj_p_.p);
        k_fill<double><<<blocks_for(cnt), kThreads, 0, st_>>>(grad_c_.p, (long)cnt, 0.0);
        LAUNCH_CHECK();
        if ((int)gbeta_.n < d.nsteps) gbeta_.alloc(d.nsteps);
        for (int seg = nseg - 1; seg >= 0; --seg) {
            const int t0 = seg * K, t1 = std::min(d.nsteps, t0 + K);
            if (seg != nseg - 1) {
                load_ckpt(*ckpts_[seg]);
                for (int t = t0; t < t1; ++t) step_recorded(d.dt, *recs_[t - t0]);
            }
            for (int t = t1 - 1; t >= t0; --t) adjoint_step(d.dt, *recs_[t - t0], d, t);
        }
        std::vector<double> per_step(d.nsteps);
        HIPT(hipMemcpyAsync(per_step.data(), gbeta_.p, d.nsteps * sizeof(double),
                            hipMemcpyDeviceToHost, st_));
        sync();
        double gbeta = 0.0;
        for (int t = d.nsteps - 1; t >= 0; --t) gbeta += per_step[t];  // reverse-sweep order
        return gbeta;
    }

    double rollout(const aether_rollout_desc &d, const double *ctrl, double beta,
                   const double *target, double *grad_c, double *grad_beta) override {
        const aether_fluid_params saved = params;
        tape_valid_ = false;
        double loss = 0.0;
        try {
            const size_t cnt = rollout_setup(d, ctrl, beta);
            if (!target_.p) target_.alloc(n_);
            HIPT(hipMemcpyAsync(target_.p, target, n_ * sizeof(double), hipMemcpyHostToDevice,
                                st_));
            const bool want = grad_c || grad_beta;
            if (want) ensure_adjoint_buffers();
            rollout_forward_sweep(d, want);
            // loss and its adjoint 2 (rho_T - target), the density seed
            const unsigned lb = partial_blocks(n_, partA_);
            k_loss<T><<<lb, kThreads, 0, st_>>>(cur_[3], target_.p, (long)n_,
                                                want ? adj_ar_.p : nullptr, partA_.p);
            LAUNCH_CHECK();
            k_partials_final<<<1, kThreads, 0, st_>>>(partA_.p, (int)lb, 0, scalar_.p);
            LAUNCH_CHECK();
            loss = read_scalar();
            if (want) {
                for (int a = 0; a < 3; ++a) zero(adj_a_[a].p);
                zero(adj_at_.p);
                const double gbeta = rollout_reverse_sweep(d, cnt);
                if (grad_c)
                    HIPT(hipMemcpyAsync(grad_c, grad_c_.p, cnt * sizeof(double),
                                        hipMemcpyDeviceToHost, st_));
                if (grad_beta) *grad_beta = gbeta;
            }
            load_ckpt(*ckpts_[0]);  // the solver's state is left as it was
            sync();
        } catch (...) {
            params = saved;
            patch_on_ = false;
            throw;
        }
        params = saved;
        patch_on_ = false;
        return loss;
    }

    // Tape form: the forward sweep keeps what the reverse sweep needs (the
    // checkpoints, the recorded last segment, the controls and beta) until
    // rollout_backward consumes it.
    void rollout_forward(const aether_rollout_desc &d, const double *ctrl, double beta,
                         double *final_state) override {
        const aether_fluid_params saved = params;
        tape_valid_ = false;
        try {
            tape_cnt_ = rollout_setup(d, ctrl, beta);
            ensure_adjoint_buffers();
            rollout_forward_sweep(d, true);
            if (final_state)
                for (int a = 0; a < 5; ++a) download_to(cur_[a], final_state + (size_t)a * n_);
            load_ckpt(*ckpts_[0]);
            sync();
        } catch (...) {
            params = saved;
            patch_on_ = false;
            throw;
        }
        tape_desc_ = d;
        tape_params_ = params;
        tape_valid_ = true;
        params = saved;
        patch_on_ = false;
    }

    void rollout_backward(const double *lam_final, double *grad_c, double *grad_beta,
                          double *grad_init) override {
        if (!tape_valid_)
            raise(AETHER_EINVAL, "rollout_backward needs a rollout_forward before it (each "
                                 "forward is consumed by one backward)");
        const aether_fluid_params saved = params;
        tape_valid_ = false;
        try {
            params = tape_params_;
            patch_on_ = true;
            T *seed[5] = {adj_a_[0].p, adj_a_[1].p, adj_a_[2].p, adj_ar_.p, adj_at_.p};
            for (int a = 0; a < 5; ++a) {
                if (lam_final)
                    upload_to(lam_final + (size_t)a * n_, seed[a]);
                else
                    zero(seed[a]);
            }
            const double gbeta = rollout_reverse_sweep(tape_desc_, tape_cnt_);
            if (grad_c)
                HIPT(hipMemcpyAsync(grad_c, grad_c_.p, tape_cnt_ * sizeof(double),
                                    hipMemcpyDeviceToHost, st_));
            if (grad_beta) *grad_beta = gbeta;
            if (grad_init) {
                const T *adj[5] = {adj_a_[0].p, adj_a_[1].p, adj_a_[2].p, adj_ar_.p, adj_at_.p};
                for (int a = 0; a < 5; ++a) download_to(adj[a], grad_init + (size_t)a * n_);
            }
            load_ckpt(*ckpts_[0]);
            sync();
        } catch (...) {
            params = saved;
            patch_on_ = false;
            throw;
        }
        params = saved;
        patch_on_ = false;
    }

    const std::vector<int> &last_run_iters() const override { return last_run_iters_; }

    int take_nonfinite() override {
        int v = 0;
        HIPT(hipMemcpyAsync(&v, nonfinite_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
        k_fill<int><<<1, kThreads, 0, st_>>>(nonfinite_.p, 1, 0);
        LAUNCH_CHECK();
        sync();
        return v;
    }

   private:
    Dims g_;
    double h_;
    hipStream_t st_;
    size_t n_;
    dim3 cell_blk_, cell_grid_, red_grid_;
    int zchunk_, nb_;
    Tiling tl_;
    int zu_ = 4;
    int tuned_form_ = 0;  // the autotuner's AETHER_CG_AUTO choice (0: none)
    int tuned_zu_ = 0, tuned_vzu_ = 0;  // its tilings (0: none)
    int zu_default_ = 4;  // z cells per thread chosen at creation
    int trial_form_ = 0;  // layout forced while the autotuner times a candidate
    int form_ = AETHER_CG_SCALAR;  // layout of the solve being enqueued
    VTiling tlv_ = {0, 0, 0, 0, 0, 1, 1};  // four-wide CG tiling (nx % 4 == 0)
    int vzu_ = 4;
    Tiling tlf_;          // one-launch CG, one cell per thread: 64 x 4 x kFusedKZ tiling
    VTiling tl4_ = {0, 0, 0, 0, 0, 1, 1};  // one-launch four-wide CG tiling (nb 0: unusable)
    DevBuf<T> r2_;                 // second residual buffer of the one-launch CG
    DevBuf<T> p2_;                 // second pressure buffer of the one-launch SOR sweep
    VTiling tls_ = {0, 0, 0, 0, 0, 1, 1};  // its tiling (2 planes per thread; nb 0: unusable)
    DevBuf<T> w_[2];               // A r of the four-wide one-launch CG (ping-pong)
    DevBuf<T> rring_[kCgRingMax - 2];  // its residual ring past r_ and r2_
    DevBuf<double> partE_, partF_;  // its delta partials
    bool ir64_ = false;           // mixed-precision projection (AETHER_FP32_IR64)
    // rollout / adjoint state
    bool patch_on_ = false;
    int3 patch_lo_ = {0, 0, 0}, patch_n_ = {0, 0, 0};
    DevBuf<double> ctrl_, grad_c_, target_, adj_part_;
    std::vector<std::unique_ptr<Ckpt>> ckpts_;
    std::vector<std::unique_ptr<StepRec>> recs_;
    DevBuf<T> adj_a_[3], adj_ar_, adj_at_, adj_b_[2], adj_c_[3], adj_d_[3], adj_tu_[3],
        adj_lw_[3], adj_lg_[3], adj_p_;
    DevBuf<fixacc_t> acc_src_, acc_fwd_;  // fixed-point scatter accumulators of the adjoint
    DevBuf<FixScale> fix_;
    DevBuf<double> gbeta_;  // per-step d loss / d beta of the reverse sweep
    // the tape between rollout_forward and rollout_backward
    bool tape_valid_ = false;
    aether_rollout_desc tape_desc_{};
    aether_fluid_params tape_params_{};
    size_t tape_cnt_ = 0;
    DevBuf<double> p64_, b64_, partD64_;
    double rel_override_ = -1.0;  // inner-solve tolerance while refining
    int ir_outer_ = 0, ir_inner_ = 0;
    double ir_div_ = 0.0;
    std::vector<int> ir_log_;
    static constexpr int kCgChunk = 8;  // CG iterations per host look-ahead chunk
    int solve_seq_ = 0;        // id of the current pressure solve
    int *done_host_ = nullptr; // mapped coherent host words: [0] the device sets to the id
                               // of a solve a CG launch stopped; [1..2] (one 64-bit word)
                               // to (id << 32) | iterations of the last finished solve
    int *done_dev_ = nullptr;
    std::vector<hipEvent_t> chunk_events_;
    DevBuf<T> f_[5], alt_[5];
    T *cur_[5], *nxt_[5];  // u, v, w, density, temperature (current / scratch)
    DevBuf<T> fwd_[2], curl_, p_, b_, r_, d_[2];
    DevBuf<double> partA_, partB_, partC_, part_tel_, scalar_, tel_, tel_ring_, staging_;
    DevBuf<double> tot_;  // pre-summed r.r and b.b of the wide four-wide path
    DevBuf<CgState> cg_;
    DevBuf<int> iters_log_, iters_dev_;
    DevBuf<int> nonfinite_;  // set when a telemetry reduction met a non-finite value
    int last_kind_ = AETHER_SOLVER_CG;
    int last_sweeps_ = 0;
    std::vector<int> last_run_iters_;
};

}  // namespace
}  // namespace aether
