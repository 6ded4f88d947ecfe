// batch.hpp -- many smoke grids advanced together, one launch per stage.
// Included by batch.hip only.
#pragma once
#include <deque>

#include "host_common.hpp"
#include "cg_vec.hpp"
#include "batch_kernels.hpp"

namespace aether {
namespace {

// ---------------------------------------------------------------------------
// Batched smoke grids: every stage is one launch over the tiles of all
// instances (batch_kernels.hpp).  Instances share dt, the solver knobs and
// the dissipation; each has its own grid, h, forcing and sphere source.
// ---------------------------------------------------------------------------
constexpr int kBatchZU = 8;  // z cells per thread of the batched CG tiling

template <typename T>
class Batch final : public BatchBase {
    struct Inst {
        Dims g;
        double h;
        size_t n;
        int gx, gy, gz, nb;
        // four-wide CG tilings (nx % 4 == 0) with 2 and 4 z planes per thread
        VTiling vtz[2] = {{0, 0, 0, 0, 0, 1, 1}, {0, 0, 0, 0, 0, 1, 1}};
        aether_fluid_instance_desc d;
        DevBuf<T> f[2][5], fwd0, fwd1, curl, p, b, r, dd[2];
        DevBuf<double> pa, pb, pc, ptel, tel, staging;
        // one-launch CG (allocated on first use): ring[1..7] beside r, w pair,
        // delta partials
        DevBuf<T> ring[kCgRingMax - 1], w[2];
        DevBuf<double> pe, pf;
        DevBuf<CgState> st;
        DevBuf<int> iters;
    };

   public:
    Batch(const aether_fluid_instance_desc *desc, int count, const aether_fluid_params &shared,
          hipStream_t stream)
        : params_(shared), st_(stream) {
        std::vector<BItem> cells, cg, cg4[2];
        std::vector<int2> src;
        // four-wide CG: FP32 and every nx % 4 == 0
        x4_ok_ = sizeof(T) == 4;
        // (its kernels store write-through, which addresses arrays under 2 GiB)
        for (int q = 0; q < count; ++q)
            x4_ok_ = x4_ok_ && desc[q].nx % 4 == 0 &&
                     (size_t)desc[q].nx * desc[q].ny * desc[q].nz * 4 < (size_t(1) << 31);
        for (int q = 0; q < count; ++q) {
            auto in = std::make_unique<Inst>();
            const aether_fluid_instance_desc &d = desc[q];
            in->d = d;
            in->g.nx = d.nx;
            in->g.ny = d.ny;
            in->g.nz = d.nz;
            in->g.sxy = d.nx * d.ny;
            in->h = d.h > 0.0 ? d.h : 1.0 / (double)std::max(d.nx, std::max(d.ny, d.nz));
            in->n = (size_t)d.nx * d.ny * d.nz;
            in->gx = (d.nx + kBX - 1) / kBX;
            in->gy = (d.ny + kBY - 1) / kBY;
            in->gz = (d.nz + kBatchZU - 1) / kBatchZU;
            in->nb = in->gx * in->gy * in->gz;
            if (x4_ok_) {  // as the single-grid four-wide tilings
                for (int zi = 0; zi < 2; ++zi) {
                    VTiling &v = in->vtz[zi];
                    const int vz = zi == 0 ? 2 : 4;
                    v.txv = std::min(d.nx / 4, 64);
                    v.ty = kThreads / v.txv;
                    v.gx = (d.nx / 4 + v.txv - 1) / v.txv;
                    v.gy = (d.ny + v.ty - 1) / v.ty;
                    v.gz = (d.nz + vz - 1) / vz;
                    v.nb = v.gx * v.gy * v.gz;
                    v.remap = (v.nb % 8 == 0) ? 1 : 0;
                    wide_[zi] = wide_[zi] || std::max(in->nb, v.nb) > 4 * kThreads;
                }
            }
            const int npart = std::max(in->nb, std::max(in->vtz[0].nb, in->vtz[1].nb));
            for (int a = 0; a < 2; ++a)
                for (int f = 0; f < 5; ++f) in->f[a][f].alloc(in->n);
            for (DevBuf<T> *b : {&in->fwd0, &in->fwd1, &in->curl, &in->p, &in->b, &in->r,
                                 &in->dd[0], &in->dd[1]})
                b->alloc(in->n);
            in->pa.alloc(npart);
            in->pb.alloc(npart);
            in->pc.alloc(npart);
            in->ptel.alloc((size_t)in->nb * kTelParts);
            in->st.alloc(1);
            for (int a = 0; a < 2; ++a)
                for (int f = 0; f < 5; ++f) fill(in->f[a][f].p, in->n);
            for (DevBuf<T> *b : {&in->fwd0, &in->fwd1, &in->curl, &in->p, &in->b, &in->r,
                                 &in->dd[0], &in->dd[1]})
                fill(b->p, in->n);
            HIPT(hipMemsetAsync(in->st.p, 0, sizeof(CgState), st_));
            // cell work items: one per (64 x 4 columns, z plane)
            for (int k = 0; k < d.nz; ++k)
                for (int by = 0; by < in->gy; ++by)
                    for (int bx = 0; bx < in->gx; ++bx) cells.push_back({q, bx, by, k, 0, {0, 0, 0}});
            // CG tiles, numbered like the single-grid launch (same XCD-slab
            // order), so each tile's partial sum lands in the same slot
            const bool remap = in->nb % 8 == 0;
            for (int b = 0; b < in->nb; ++b) {
                const int t = remap ? (b & 7) * (in->nb >> 3) + (b >> 3) : b;
                const int bx = t % in->gx, rest = t / in->gx;
                cg.push_back({q, bx, rest % in->gy, rest / in->gy, b, {0, 0, 0}});
            }
            if (x4_ok_)
                for (int zi = 0; zi < 2; ++zi) {
                    const VTiling &v = in->vtz[zi];
                    for (int b = 0; b < v.nb; ++b) {
                        const int t = v.remap ? (b & 7) * (v.nb >> 3) + (b >> 3) : b;
                        const int bx = t % v.gx, rest = t / v.gx;
                        cg4[zi].push_back({q, bx, rest % v.gy, rest / v.gy, b, {0, 0, 0}});
                    }
                }
            // sphere source box (interior cells)
            if (d.source_radius > 0.0) {
                const int nn[3] = {d.nx, d.ny, d.nz};
                int lo[3], hi[3];
                bool empty = false;
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::max(1, (int)std::floor(d.source_center[a] - d.source_radius));
                    hi[a] = std::min(nn[a] - 2, (int)std::ceil(d.source_center[a] + d.source_radius));
                    if (hi[a] < lo[a]) empty = true;
                }
                if (!empty) {
                    const int cnt = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
                    for (int a = 0; a < cnt; a += kThreads) src.push_back(make_int2(q, a));
                    for (int a = 0; a < 3; ++a) {
                        box_lo_.push_back(lo[a]);
                        box_n_.push_back(hi[a] - lo[a] + 1);
                    }
                } else {
                    for (int a = 0; a < 3; ++a) {
                        box_lo_.push_back(1);
                        box_n_.push_back(0);
                    }
                }
            } else {
                for (int a = 0; a < 3; ++a) {
                    box_lo_.push_back(1);
                    box_n_.push_back(0);
                }
            }
            inst_.push_back(std::move(in));
        }
        ncell_ = (int)cells.size();
        ncg_ = (int)cg.size();
        nsrc_ = (int)src.size();
        cell_items_.alloc(ncell_);
        cg_items_.alloc(ncg_);
        for (int zi = 0; zi < 2; ++zi) {
            ncg4_[zi] = (int)cg4[zi].size();
            if (ncg4_[zi]) {
                cg4_items_[zi].alloc(ncg4_[zi]);
                HIPT(hipMemcpyAsync(cg4_items_[zi].p, cg4[zi].data(),
                                    cg4[zi].size() * sizeof(BItem), hipMemcpyHostToDevice, st_));
            }
        }
        HIPT(hipMemcpyAsync(cell_items_.p, cells.data(), cells.size() * sizeof(BItem),
                            hipMemcpyHostToDevice, st_));
        HIPT(hipMemcpyAsync(cg_items_.p, cg.data(), cg.size() * sizeof(BItem),
                            hipMemcpyHostToDevice, st_));
        if (nsrc_) {
            src_items_.alloc(nsrc_);
            HIPT(hipMemcpyAsync(src_items_.p, src.data(), src.size() * sizeof(int2),
                                hipMemcpyHostToDevice, st_));
        }
        table_.alloc(count);
        // per instance: id of the last solve a CG launch stopped; coherent, so
        // the device's writes are seen by the host's pacing polls without a flush
        HIPT(hipHostMalloc((void **)&done_host_, sizeof(int) * count,
                           hipHostMallocMapped | hipHostMallocCoherent));
        for (int q = 0; q < count; ++q) done_host_[q] = 0;
        HIPT(hipHostGetDevicePointer((void **)&done_dev_, done_host_, 0));
        HIPT(hipStreamSynchronize(st_));
    }

    ~Batch() override {
        (void)hipStreamSynchronize(st_);
        for (auto e : chunk_events_) (void)hipEventDestroy(e);
        if (done_host_) (void)hipHostFree(done_host_);
    }

    int count() const override { return (int)inst_.size(); }
    const std::vector<int> &iters() const override { return iters_; }

    void fill(T *x, size_t n) {
        const unsigned b = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
        k_fill<T><<<std::max(b, 1u), kThreads, 0, st_>>>(x, (long)n, T(0));
        LAUNCH_CHECK();
    }

    T *field(Inst &in, int field) {
        const int slots[5] = {sl_.u, sl_.v, sl_.w, sl_.rho, sl_.t};
        if (field >= 0 && field < 5) return in.f[slots[field]][field].p;
        if (field == AETHER_FIELD_PRESSURE) return in.p.p;
        if (field == AETHER_FIELD_DIVERGENCE) return in.b.p;
        raise(AETHER_EINVAL, "unknown field id " + std::to_string(field));
    }

    Inst &inst(int q) {
        if (q < 0 || q >= count()) raise(AETHER_EINVAL, "instance index out of range");
        return *inst_[q];
    }

    void upload(int q, int f, const double *src, size_t n) override {
        Inst &in = inst(q);
        if (n != in.n) raise(AETHER_EINVAL, "upload size does not match the instance grid");
        T *dst = field(in, f);
        if (!in.staging.p) in.staging.alloc(in.n);
        HIPT(hipMemcpyAsync(in.staging.p, src, n * sizeof(double), hipMemcpyHostToDevice, st_));
        k_convert<double, T><<<4096, kThreads, 0, st_>>>(in.staging.p, (long)n, dst);
        LAUNCH_CHECK();
        HIPT(hipStreamSynchronize(st_));
    }

    void download(int q, int f, double *dst, size_t n) override {
        Inst &in = inst(q);
        if (n != in.n) raise(AETHER_EINVAL, "download size does not match the instance grid");
        T *src = field(in, f);
        if (!in.staging.p) in.staging.alloc(in.n);
        k_convert<T, double><<<4096, kThreads, 0, st_>>>(src, (long)n, in.staging.p);
        LAUNCH_CHECK();
        HIPT(hipMemcpyAsync(dst, in.staging.p, n * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPT(hipStreamSynchronize(st_));
    }

    void write_table() {
        std::vector<BInst<T>> tab(count());
        for (int q = 0; q < count(); ++q) {
            Inst &in = *inst_[q];
            BInst<T> &t = tab[q];
            std::memset(&t, 0, sizeof(t));
            t.g = in.g;
            t.h = in.h;
            for (int a = 0; a < 2; ++a)
                for (int f = 0; f < 5; ++f) t.f[a][f] = in.f[a][f].p;
            t.fwd0 = in.fwd0.p;
            t.fwd1 = in.fwd1.p;
            t.curl = in.curl.p;
            t.p = in.p.p;
            t.b = in.b.p;
            t.r = in.r.p;
            t.d[0] = in.dd[0].p;
            t.d[1] = in.dd[1].p;
            t.pa = in.pa.p;
            t.pb = in.pb.p;
            t.pc = in.pc.p;
            t.ptel = in.ptel.p;
            t.st = in.st.p;
            t.nb = in.nb;
            t.eps = in.d.vorticity_eps;
            t.beta = in.d.buoyancy_beta;
            t.alpha = in.d.gravity_alpha;
            for (int a = 0; a < 3; ++a) {
                t.sc[a] = in.d.source_center[a];
                t.slo[a] = box_lo_[3 * q + a];
                t.sn[a] = box_n_[3 * q + a];
            }
            t.sr2 = in.d.source_radius * in.d.source_radius;
            t.srho = in.d.source_density;
            t.stemp = in.d.source_temperature;
            t.tel = in.tel.p;
            t.iters = in.iters.p;
            t.vt = in.vtz[vzi()];
            t.nbv = t.vt.nb;
            t.ring[0] = in.r.p;
            for (int q = 1; q < kCgRingMax; ++q) t.ring[q] = in.ring[q - 1].p;
            t.w[0] = in.w[0].p;
            t.w[1] = in.w[1].p;
            t.pe = in.pe.p;
            t.pf = in.pf.p;
            t.vt4 = in.vtz[1];
            t.nb4 = in.vtz[1].nb;
        }
        HIPT(hipMemcpyAsync(table_.p, tab.data(), tab.size() * sizeof(BInst<T>),
                            hipMemcpyHostToDevice, st_));
        HIPT(hipStreamSynchronize(st_));  // tab is pageable and ends with this scope
    }

    void advect1(int f, double dt, double diss) {
        kb_sl_forward<T, 1><<<ncell_, dim3(kBX, kBY), 0, st_>>>(table_.p, cell_items_.p, dt, sl_, f);
        LAUNCH_CHECK();
        kb_mac_correct<T, 1><<<ncell_, dim3(kBX, kBY), 0, st_>>>(table_.p, cell_items_.p, dt, sl_,
                                                                f, diss, 0.0, 0);
        LAUNCH_CHECK();
        int *slot[5] = {&sl_.u, &sl_.v, &sl_.w, &sl_.rho, &sl_.t};
        *slot[f] ^= 1;
    }

    // The four-wide tiling in use (0: 2 z planes per thread, 1: 4): the
    // autotuner's trial, else cg_z_cells, else the autotuner's choice, else 4.
    int vzi() const {
        int vz = 4;
        if (trial_vz_)
            vz = trial_vz_;
        else if (params_.cg_z_cells == 2 || params_.cg_z_cells == 4)
            vz = params_.cg_z_cells;
        else if (tuned_vz_)
            vz = tuned_vz_;
        return vz == 2 ? 0 : 1;
    }

    // The batch's CG layout for the current knobs: AETHER_CG_X4 (four-wide,
    // two launches per iteration over every instance), AETHER_CG_FUSED (one
    // launch per iteration, solved group by group) or AETHER_CG_SCALAR.
    int form() const {
        if (trial_form_) return trial_form_;
        switch (params_.cg_form) {
        case AETHER_CG_SCALAR:
            return AETHER_CG_SCALAR;
        case AETHER_CG_X4:
            if (!x4_ok_) raise(AETHER_EINVAL, "AETHER_CG_X4 needs FP32 and nx % 4 == 0 in every instance");
            return AETHER_CG_X4;
        case AETHER_CG_FUSED:
            if (!x4_ok_)
                raise(AETHER_EINVAL, "AETHER_CG_FUSED for batches needs FP32 and nx % 4 == 0 "
                                     "in every instance");
            return AETHER_CG_FUSED;
        default:
            if (tuned_form_) return tuned_form_;
            return x4_ok_ ? AETHER_CG_X4 : AETHER_CG_SCALAR;
        }
    }
    bool use_x4() const { return form() == AETHER_CG_X4; }

    // The one-launch CG's buffers, and its groups: instances whose launches
    // share a kernel (the same tiling class and partial-sum width, for the
    // first launch and for the others), packed in instance order into groups
    // of at most kOneLaunchGroupCells cells.  Measured on the 64-grid workload
    // (ms per step at caps of 3.2 / 4.8 / 6.4 / 9.6 / 16 M cells and no cap:
    // 39.5 / 37.3 / 36.0 / 36.6 / 38.4 / 38.2): larger launches pay the launch
    // floor less often until a group's arrays (~50 bytes per cell) spill out
    // of the 256 MB Infinity Cache between the launches that reuse them.
    static constexpr size_t kOneLaunchGroupCells = 6400000;
    void ensure_fused() {
        if (fused_ready_) return;
        if (!x4_ok_)
            raise(AETHER_EINVAL, "AETHER_CG_FUSED for batches needs FP32 and nx % 4 == 0 in "
                                 "every instance");
        for (auto &in : inst_) {
            for (DevBuf<T> &b : in->ring) {
                b.alloc(in->n);
                fill(b.p, in->n);
            }
            for (DevBuf<T> &b : in->w) {
                b.alloc(in->n);
                fill(b.p, in->n);
            }
            in->pe.alloc(in->pa.n);
            in->pf.alloc(in->pa.n);
        }
        groups_.clear();
        for (int q = 0; q < count(); ++q) {
            const Inst &in = *inst_[q];
            const VTiling &v = in.vtz[1];
            const int rg = !vtiling_regular(v, in.g.ny, in.g.nz, 4) ? 0 : v.txv == 32 ? 2 : 1;
            const int c0 = kb_cgs4_class(std::max(in.nb, v.nb), rg), c1 = kb_cgs4_class(v.nb, rg);
            Group *gp = nullptr;
            for (Group &gr : groups_)
                if (gr.open && gr.cls_first == c0 && gr.cls == c1) gp = &gr;
            if (gp && gp->cells + in.n > kOneLaunchGroupCells) {
                gp->open = false;
                gp = nullptr;
            }
            if (!gp) {
                groups_.emplace_back();
                gp = &groups_.back();
                gp->cls_first = c0;
                gp->cls = c1;
            }
            gp->insts.push_back(q);
            gp->cells += in.n;
        }
        std::deque<std::vector<BItem>> host;  // kept until the copies below have completed
        for (Group &gr : groups_) {
            std::vector<BItem> &items = host.emplace_back();
            for (int q : gr.insts) {
                const VTiling &v = inst_[q]->vtz[1];
                for (int b = 0; b < v.nb; ++b) {  // the single-grid launch order
                    const int t = v.remap ? (b & 7) * (v.nb >> 3) + (b >> 3) : b;
                    const int bx = t % v.gx, rest = t / v.gx;
                    items.push_back({q, bx, rest % v.gy, rest / v.gy, b, {0, 0, 0}});
                }
            }
            gr.nitems = (int)items.size();
            gr.items.alloc(items.size());
            HIPT(hipMemcpyAsync(gr.items.p, items.data(), items.size() * sizeof(BItem),
                                hipMemcpyHostToDevice, st_));
        }
        HIPT(hipStreamSynchronize(st_));
        fused_ready_ = true;
    }

    // Times each candidate (four-wide with 2 and 4 z planes per thread when
    // every nx % 4 == 0, one cell per thread with 8, and the one-launch form
    // solved group by group) and keeps the fastest as the choice for cg_form
    // AETHER_CG_AUTO and cg_z_cells 0; every instance's pressure and
    // divergence are saved and restored.  iters > 0: `iters` iterations on
    // synthetic right-hand sides with no tolerance stop.  iters 0: whole
    // solves, from zero pressure, of the last step's right-hand sides at the
    // batch's tolerance and iteration cap, so that instances converging at
    // different iterations weigh in as in the batch's own steps.  us[4]: time
    // per iteration of the whole batch (iters 0: per iteration of the
    // instance that ran longest), in that order (-1 when not applicable).
    void autotune(int iters, double *us, int *form_out, int *z_out) override {
        if (iters < 0) raise(AETHER_EINVAL, "iters must be >= 0");
        // iters 0: whole solves of the last step's right-hand side
        const bool real = iters == 0;
        if (real && !rhs_ready_)
            raise(AETHER_EINVAL, "autotune with iters 0 solves the last step's right-hand "
                                 "side: run a step first");
        for (auto &in : inst_) {  // finalize logs iteration 0's count
            if (in->iters.n < 1) in->iters.alloc(1);
            if (in->tel.n < 8) in->tel.alloc(8);
        }
        std::vector<std::unique_ptr<DevBuf<T>>> saved;
        for (auto &in : inst_) {
            for (DevBuf<T> *src : {&in->p, &in->b}) {
                saved.push_back(std::make_unique<DevBuf<T>>(in->n));
                HIPT(hipMemcpyAsync(saved.back()->p, src->p, in->n * sizeof(T),
                                    hipMemcpyDeviceToDevice, st_));
            }
            if (!real) {
                k_fill_hash<T><<<4096, kThreads, 0, st_>>>(in->b.p, (long)in->n);
                LAUNCH_CHECK();
            }
        }
        const aether_fluid_params keep = params_;
        if (!real) {
            params_.cg_rel_tol = 0.0;  // absolute guards only: each candidate runs `iters`
            params_.solver_iters = iters;
        }
        std::vector<int> its(count());
        hipEvent_t e0, e1;
        HIPT(hipEventCreate(&e0));
        HIPT(hipEventCreate(&e1));
        struct Cand {
            int form, z;
        };
        const Cand cands[4] = {{AETHER_CG_X4, 2},
                               {AETHER_CG_X4, 4},
                               {AETHER_CG_SCALAR, kBatchZU},
                               {AETHER_CG_FUSED, 4}};
        double best = 1e300;
        Cand pick = {AETHER_CG_SCALAR, kBatchZU};
        auto cleanup = [&] {
            trial_form_ = 0;
            trial_vz_ = 0;
            params_ = keep;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        };
        try {
            for (int a = 0; a < 4; ++a) {
                if (us) us[a] = -1.0;
                if (cands[a].form != AETHER_CG_SCALAR && !x4_ok_) continue;
                if (cands[a].form == AETHER_CG_FUSED) ensure_fused();
                trial_form_ = cands[a].form;
                trial_vz_ = cands[a].form == AETHER_CG_X4 ? cands[a].z : 0;
                write_table();  // the table holds the four-wide tiling in use
                float ms = 0.0f;
                for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
                    for (auto &in : inst_) fill(in->p.p, in->n);  // from p = 0
                    kb_div_cginit<T, kBatchZU><<<ncg_, dim3(kBX, kBY), 0, st_>>>(
                        table_.p, cg_items_.p, sl_, 0);
                    LAUNCH_CHECK();
                    HIPT(hipEventRecord(e0, st_));
                    cg_solve(0);
                    HIPT(hipEventRecord(e1, st_));
                    HIPT(hipEventSynchronize(e1));
                    HIPT(hipEventElapsedTime(&ms, e0, e1));
                }
                int nit = iters;
                if (real) {  // per iteration of the instance that ran longest
                    nit = 1;
                    for (int q = 0; q < count(); ++q) {
                        HIPT(hipMemcpy(&its[q], inst_[q]->iters.p, sizeof(int),
                                       hipMemcpyDeviceToHost));
                        nit = std::max(nit, its[q]);
                    }
                }
                const double t = 1e3 * (double)ms / nit;
                if (us) us[a] = t;
                if (t < best) {
                    best = t;
                    pick = cands[a];
                }
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        tuned_form_ = pick.form;
        if (pick.form == AETHER_CG_X4) tuned_vz_ = pick.z;
        write_table();
        size_t k = 0;
        for (auto &in : inst_)
            for (DevBuf<T> *dst : {&in->p, &in->b}) {
                HIPT(hipMemcpyAsync(dst->p, saved[k]->p, in->n * sizeof(T),
                                    hipMemcpyDeviceToDevice, st_));
                ++k;
            }
        HIPT(hipStreamSynchronize(st_));
        if (form_out) *form_out = pick.form;
        if (z_out) *z_out = pick.z;
    }

    void launch_cg4(int it, double rel2, int solve_id) {
        if constexpr (sizeof(T) == 4) {
            // p is brought up to date on odd iterations (kb_cg_update4)
            const bool odd = it & 1;
            const int zi = vzi();
            const dim3 g(ncg4_[zi]), b(kThreads);
            const BItem *items = cg4_items_[zi].p;
#define KB_DIR4(VZ_, NQ_)                                                                    \
    do {                                                                                     \
        if (it == 0)                                                                         \
            kb_cg_dir4<VZ_, true, NQ_, false><<<g, b, 0, st_>>>(table_.p, items, it, rel2,   \
                                                                done_dev_, solve_id);        \
        else if (odd)                                                                        \
            kb_cg_dir4<VZ_, false, NQ_, true><<<g, b, 0, st_>>>(table_.p, items, it, rel2,   \
                                                                done_dev_, solve_id);        \
        else                                                                                 \
            kb_cg_dir4<VZ_, false, NQ_, false><<<g, b, 0, st_>>>(table_.p, items, it, rel2,  \
                                                                 done_dev_, solve_id);       \
        LAUNCH_CHECK();                                                                      \
        if (odd)                                                                             \
            kb_cg_update4<VZ_, NQ_, kPBoth><<<g, b, 0, st_>>>(table_.p, items, it,           \
                                                              done_dev_, solve_id);          \
        else                                                                                 \
            kb_cg_update4<VZ_, NQ_, kPSkip><<<g, b, 0, st_>>>(table_.p, items, it,           \
                                                              done_dev_, solve_id);          \
        LAUNCH_CHECK();                                                                      \
    } while (0)
            if (zi == 0) {
                if (wide_[0])
                    KB_DIR4(2, 16);
                else
                    KB_DIR4(2, 4);
            } else {
                if (wide_[1])
                    KB_DIR4(4, 16);
                else
                    KB_DIR4(4, 4);
            }
#undef KB_DIR4
        }
    }

    // graph: the step is being captured, so every iteration is queued (no
    // host look-ahead) and the step index comes from the device.
    void cg_solve(int step, bool graph = false) {
        const int max_iters = params_.solver_iters;
        const int fm = form();
        const bool x4 = fm == AETHER_CG_X4;
        const int zc = params_.cg_z_cells;
        if (!trial_form_ && zc) {
            if (fm == AETHER_CG_X4 && zc != 2 && zc != 4)
                raise(AETHER_EINVAL, "the batch's four-wide CG takes cg_z_cells 2 or 4");
            if (fm == AETHER_CG_FUSED && zc != 4)
                raise(AETHER_EINVAL, "the batch's one-launch CG runs 4 z planes per thread "
                                     "(cg_z_cells 0 or 4)");
            if (fm == AETHER_CG_SCALAR && zc != kBatchZU)
                raise(AETHER_EINVAL, "the batch's one-cell-per-thread CG runs 8 z cells per "
                                     "thread (cg_z_cells 0 or 8)");
        }
        if (fm == AETHER_CG_FUSED) {
            cg_solve_fused(step, graph);
            return;
        }
        const double rel2 = params_.cg_rel_tol > 0 ? params_.cg_rel_tol * params_.cg_rel_tol : 0.0;
        const int solve_id = ++solve_seq_;
        const int chunk = 8;
        const int nchunks = (max_iters + chunk - 1) / chunk;
        while ((int)chunk_events_.size() < nchunks) {
            hipEvent_t e;
            // pacing only (as the single-grid look-ahead events)
            HIPT(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
            chunk_events_.push_back(e);
        }
        const dim3 blk(kBX, kBY);
        for (int c = 0; c < nchunks; ++c) {
            if (c >= 2 && !graph) {
                HIPT(hipEventSynchronize(chunk_events_[c - 2]));
                bool all = true;
                for (int q = 0; q < count() && all; ++q)
                    all = __atomic_load_n(done_host_ + q, __ATOMIC_ACQUIRE) == solve_id;
                if (all) break;
            }
            for (int it = c * chunk; it < std::min(max_iters, (c + 1) * chunk); ++it) {
                if (x4) {
                    launch_cg4(it, rel2, solve_id);
                    continue;
                }
                if (it == 0)
                    kb_cg_dir<T, kBatchZU, true><<<ncg_, blk, 0, st_>>>(
                        table_.p, cg_items_.p, it, rel2, done_dev_, solve_id);
                else
                    kb_cg_dir<T, kBatchZU, false><<<ncg_, blk, 0, st_>>>(
                        table_.p, cg_items_.p, it, rel2, done_dev_, solve_id);
                LAUNCH_CHECK();
                kb_cg_update<T, kBatchZU><<<ncg_, blk, 0, st_>>>(table_.p, cg_items_.p, it,
                                                                 done_dev_, solve_id);
                LAUNCH_CHECK();
            }
            if (!graph) HIPT(hipEventRecord(chunk_events_[c], st_));
        }
        if constexpr (sizeof(T) == 4) {
            if (x4 && (max_iters & 1)) {  // an odd cap leaves the last alpha d pending
                const int zi = vzi();
                if (zi == 0)
                    kb_cg_p_tail<2><<<ncg4_[0], kThreads, 0, st_>>>(table_.p, cg4_items_[0].p,
                                                                   max_iters - 1);
                else
                    kb_cg_p_tail<4><<<ncg4_[1], kThreads, 0, st_>>>(table_.p, cg4_items_[1].p,
                                                                   max_iters - 1);
                LAUNCH_CHECK();
            }
        }
        kb_cg_finalize<T><<<count(), blk, 0, st_>>>(table_.p, max_iters, rel2, step, x4 ? 1 : 0,
                                                    graph ? step_dev_.p : nullptr);
        LAUNCH_CHECK();
    }

    // The one-launch CG (cg_cgs4.hpp), solved group by group: every launch of
    // a group's solve runs before the next group starts, so that the group's
    // ring, w, d and p stay in the Infinity Cache.  Within a group the
    // launches are paced as the other forms (chunks of 8, two ahead, stop
    // once every instance of the group has published this solve as finished);
    // an instance that converged first returns from each later launch before
    // loading anything.
    void cg_solve_fused(int step, bool graph) {
        if constexpr (sizeof(T) == 4) {
            ensure_fused();
            const int max_iters = params_.solver_iters;
            const double rel2 =
                params_.cg_rel_tol > 0 ? params_.cg_rel_tol * params_.cg_rel_tol : 0.0;
            const int solve_id = ++solve_seq_;
            const int chunk = 8;
            const int nchunks = (max_iters + chunk - 1) / chunk;
            while ((int)chunk_events_.size() < nchunks) {
                hipEvent_t e;
                HIPT(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence));
                chunk_events_.push_back(e);
            }
            const dim3 blk(kBX, kBY);
            if (max_iters > 0) {  // delta_0 partials and w_0 = A r_0 of every instance
                kb_cg_rar<T, kBatchZU><<<ncg_, blk, 0, st_>>>(table_.p, cg_items_.p);
                LAUNCH_CHECK();
            }
            const BInst<T> *tab = table_.p;
            int *dm = done_dev_;
            for (Group &gr : groups_) {
                const BItem *items = gr.items.p;
                for (int c = 0; c < nchunks; ++c) {
                    if (c >= 2 && !graph) {
                        HIPT(hipEventSynchronize(chunk_events_[c - 2]));
                        bool all = true;
                        for (int q : gr.insts)
                            if (__atomic_load_n(done_host_ + q, __ATOMIC_ACQUIRE) != solve_id) {
                                all = false;
                                break;
                            }
                        if (all) break;
                    }
                    for (int it = c * chunk; it < std::min(max_iters, (c + 1) * chunk); ++it) {
                        int itv = it, sid = solve_id;
                        double rt = rel2;
                        void *args[] = {&tab, &items, &itv, &rt, &dm, &sid};
                        const void *fn = kb_cgs4_kernel(it == 0, it % kCgsCycle,
                                                        it == 0 ? gr.cls_first : gr.cls);
                        HIPT(hipLaunchKernel(fn, dim3(gr.nitems), dim3(kThreads), args, 0, st_));
                    }
                    if (!graph) HIPT(hipEventRecord(chunk_events_[c], st_));
                }
            }
            // a cap that ends mid-cycle leaves directions unbuilt and p terms
            // pending (the residuals launch j read sit in ring[j % K])
            if (max_iters > 0 && max_iters % kCgsCycle != 0) {
                const int last = max_iters - 1, j0 = last - last % kCgsCycle;
                for (auto &in : inst_) {
                    CgRing pend{};
                    for (int q = 0; q <= last % kCgsCycle; ++q) {
                        const int slot = (j0 + q) % kCgsCycle;
                        pend.r[q] = slot == 0 ? in->r.p : in->ring[slot - 1].p;
                    }
                    const long n4 = (long)(in->n / 4);
                    const unsigned nbt =
                        (unsigned)std::min<long>((n4 + kThreads - 1) / kThreads, 4096);
                    cgs4_tail_launch(nbt, st_, in->st.p, last, pend,
                                     reinterpret_cast<const float4 *>(in->dd[0].p),
                                     reinterpret_cast<float4 *>(in->p.p), n4);
                    LAUNCH_CHECK();
                }
            }
            kb_cg_finalize<T><<<count(), blk, 0, st_>>>(table_.p, max_iters, rel2, step, 2,
                                                        graph ? step_dev_.p : nullptr);
            LAUNCH_CHECK();
        } else {
            raise(AETHER_EINVAL, "AETHER_CG_FUSED for batches needs FP32");
        }
    }

    // src/fluid.rs:598-611 for every instance at once.  graph: the launches
    // are being captured; the step index and time are read on the device.
    void step(double dt, double t, int s, bool graph = false) {
        const dim3 blk(kBX, kBY);
        advect1(0, dt, params_.dissipation_velocity);
        advect1(1, dt, params_.dissipation_velocity);
        advect1(2, dt, params_.dissipation_velocity);
        if (nsrc_) {
            kb_sphere<T><<<nsrc_, blk, 0, st_>>>(table_.p, src_items_.p, sl_);
            LAUNCH_CHECK();
        }
        kb_curl_mag<T><<<ncell_, blk, 0, st_>>>(table_.p, cell_items_.p, sl_);
        LAUNCH_CHECK();
        kb_confine_buoy_bc<T><<<ncell_, blk, 0, st_>>>(table_.p, cell_items_.p, dt, sl_);
        LAUNCH_CHECK();
        sl_.u ^= 1;
        sl_.v ^= 1;
        sl_.w ^= 1;
        kb_div_cginit<T, kBatchZU><<<ncg_, blk, 0, st_>>>(table_.p, cg_items_.p, sl_, 1);
        rhs_ready_ = true;
        LAUNCH_CHECK();
        cg_solve(s, graph);
        kb_gradient_bc<T><<<ncell_, blk, 0, st_>>>(table_.p, cell_items_.p, sl_);
        LAUNCH_CHECK();
        kb_sl_forward<T, 2><<<ncell_, blk, 0, st_>>>(table_.p, cell_items_.p, dt, sl_, 3);
        LAUNCH_CHECK();
        kb_mac_correct<T, 2><<<ncell_, blk, 0, st_>>>(table_.p, cell_items_.p, dt, sl_, 3,
                                                      params_.dissipation_density,
                                                      params_.dissipation_temperature, 1);
        LAUNCH_CHECK();
        sl_.rho ^= 1;
        sl_.t ^= 1;
        kb_telemetry_part<T, kBatchZU><<<ncg_, blk, 0, st_>>>(table_.p, cg_items_.p, sl_);
        LAUNCH_CHECK();
        kb_telemetry_final<T><<<count(), blk, 0, st_>>>(table_.p, s, t,
                                                        graph ? step_dev_.p : nullptr,
                                                        graph ? times_dev_.p : nullptr);
        LAUNCH_CHECK();
        if (graph) {
            kb_step_advance<<<1, 64, 0, st_>>>(step_dev_.p);
            LAUNCH_CHECK();
        }
    }

    // Captures one step into an executable graph (the host-side slot
    // bookkeeping advances as if the step had run).
    hipGraphExec_t capture_step(double dt) {
        hipGraph_t graph = nullptr;
        HIPT(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
        try {
            step(dt, 0.0, 0, true);
        } catch (...) {
            (void)hipStreamEndCapture(st_, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIPT(hipStreamEndCapture(st_, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPT(e);
        return exec;
    }

    // Steps as graph replays: the ping-pong slots of density/temperature
    // alternate per step, so two graphs (one per parity) serve every step.
    void run_graphs(double dt, const double *times, int nsteps) {
        if (times_dev_.n < (size_t)nsteps) times_dev_.alloc(nsteps);
        HIPT(hipMemcpyAsync(times_dev_.p, times, nsteps * sizeof(double), hipMemcpyHostToDevice,
                            st_));
        if (!step_dev_.p) step_dev_.alloc(1);
        k_fill<int><<<1, kThreads, 0, st_>>>(step_dev_.p, 1L, 0);
        LAUNCH_CHECK();
        HIPT(hipStreamSynchronize(st_));
        struct Execs {
            hipGraphExec_t g[2] = {nullptr, nullptr};
            ~Execs() {
                for (auto e : g)
                    if (e) (void)hipGraphExecDestroy(e);
            }
        } ex;
        const Slots start = sl_;
        ex.g[0] = capture_step(dt);
        if (nsteps > 1) ex.g[1] = capture_step(dt);
        sl_ = start;
        for (int s = 0; s < nsteps; ++s) HIPT(hipGraphLaunch(ex.g[s & 1], st_));
        if (nsteps & 1) {  // the slots the replays left current
            sl_.rho ^= 1;
            sl_.t ^= 1;
        }
        HIPT(hipStreamSynchronize(st_));
    }

    void run(double dt, const double *times, int nsteps, aether_fluid_telemetry *out) override {
        if (nsteps <= 0) return;
        for (auto &in : inst_) {
            if (in->tel.n < (size_t)nsteps * 8) in->tel.alloc((size_t)nsteps * 8);
            if (in->iters.n < (size_t)nsteps) in->iters.alloc(nsteps);
        }
        if (form() == AETHER_CG_FUSED) ensure_fused();
        write_table();
        if (params_.batch_graph)
            run_graphs(dt, times, nsteps);
        else
            for (int s = 0; s < nsteps; ++s) step(dt, times[s], s);
        const int cnt = count();
        std::vector<double> raw((size_t)nsteps * 8);
        iters_.assign((size_t)nsteps * cnt, 0);
        std::vector<int> it(nsteps);
        for (int q = 0; q < cnt; ++q) {
            HIPT(hipMemcpyAsync(raw.data(), inst_[q]->tel.p, raw.size() * sizeof(double),
                                hipMemcpyDeviceToHost, st_));
            HIPT(hipMemcpyAsync(it.data(), inst_[q]->iters.p, nsteps * sizeof(int),
                                hipMemcpyDeviceToHost, st_));
            HIPT(hipStreamSynchronize(st_));
            for (int s = 0; s < nsteps; ++s) {
                iters_[(size_t)s * cnt + q] = it[s];
                if (out) {
                    aether_fluid_telemetry &o = out[(size_t)s * cnt + q];
                    const double *r = raw.data() + (size_t)s * 8;
                    o.time_sec = r[0];
                    o.kinetic_energy = r[1];
                    o.enstrophy = r[2];
                    o.peak_vorticity = r[3];
                    o.total_mass = r[4];
                    for (int a = 0; a < 3; ++a) o.center_of_mass[a] = r[5 + a];
                }
            }
        }
    }

   private:
    aether_fluid_params params_;
    hipStream_t st_;
    std::vector<std::unique_ptr<Inst>> inst_;
    std::vector<int> box_lo_, box_n_;
    DevBuf<BInst<T>> table_;
    DevBuf<BItem> cell_items_, cg_items_, cg4_items_[2];
    DevBuf<int2> src_items_;
    int ncell_ = 0, ncg_ = 0, nsrc_ = 0, ncg4_[2] = {0, 0};
    bool x4_ok_ = false;  // FP32 and every nx % 4 == 0: four-wide CG available
    int tuned_form_ = 0;  // the autotuner's AETHER_CG_AUTO choice (0: none)
    int trial_form_ = 0;  // layout forced while the autotuner times it
    DevBuf<int> step_dev_;          // device step index of captured steps
    DevBuf<double> times_dev_;      // their time axis
    bool wide_[2] = {false, false};  // some instance has more than 1024 CG partials
    int tuned_vz_ = 0;    // the autotuner's z planes per thread of the four-wide tiling
    int trial_vz_ = 0;    // planes forced while the autotuner times them
    Slots sl_ = {0, 0, 0, 0, 0};
    int *done_host_ = nullptr, *done_dev_ = nullptr;
    int solve_seq_ = 0;
    std::vector<hipEvent_t> chunk_events_;
    std::vector<int> iters_;
    struct Group {
        std::vector<int> insts;
        size_t cells = 0;
        int cls_first = 0, cls = 0;  // kb_cgs4_class of the first launch and of the others
        bool open = true;            // still taking instances
        DevBuf<BItem> items;
        int nitems = 0;
    };
    std::deque<Group> groups_;  // (a deque: groups hold device buffers and never move)
    bool fused_ready_ = false;
    bool rhs_ready_ = false;  // b holds a step's right-hand side (autotune with iters 0)
};

}  // namespace
}  // namespace aether
