// batch_kernels.hpp -- many independent smoke grids of different sizes in one
// launch per stage.
//
// Every batched kernel reads a work list of (instance, tile) items; the
// instance table gives that grid's dimensions, buffers and forcing.  The
// per-tile work is the same device body as the single-grid kernels, so a
// batched step computes exactly what the instances would compute one at a
// time (tested bit for bit).  Each instance keeps its own CG scalars,
// partial sums and stop flag: a converged instance's tiles return at once
// while the others keep iterating.
#pragma once
#include "cg_cgs4.hpp"
#include "cg_fused.hpp"
#include "cg_vec.hpp"

namespace aether {
namespace {

struct BItem {
    int inst;       // instance index
    int bx, by, bz; // tile coordinates inside the instance
    int local;      // tile index inside the instance (its partial-sum slot)
    int pad[3];
};

template <typename T>
struct BInst {
    Dims g;
    double h;
    T *f[2][5];  // [slot][u, v, w, density, temperature]
    T *fwd0, *fwd1, *curl, *p, *b, *r, *d[2];
    double *pa, *pb, *pc, *ptel;  // this instance's partial-sum segments
    CgState *st;
    int nb;                // tiles of the instance in the CG work list
    int pad;
    double eps, beta, alpha;              // forcing
    double sc[3], sr2, srho, stemp;       // sphere source (cell-index coordinates)
    int slo[3], sn[3];                    // sphere bounding box, interior-clamped
    double *tel;                          // telemetry ring, 8 doubles per step
    int *iters;                           // CG iterations per step
    VTiling vt;                           // four-wide CG tiling (FP32, nx % 4 == 0)
    int nbv;                              // its tile count (= partial sums)
    int pad2;
    // one-launch CG (FP32, nx % 4 == 0): the residual ring (ring[0] = r), the
    // ping-pong pair of w = A r, the delta partials (ping-pong with pe/pf as
    // gamma with pb/pa), and its tiling (four z planes per thread)
    T *ring[kCgRingMax];
    T *w[2];
    double *pe, *pf;
    VTiling vt4;
    int nb4;
    int pad3;
};

struct Slots {
    int u, v, w, rho, t;  // slot (0/1) holding the current value of each field
};

template <typename T>
__device__ __forceinline__ void cell_of(const BItem &it, int &i, int &j, int &k) {
    i = it.bx * kBX + threadIdx.x;
    j = it.by * kBY + threadIdx.y;
    k = it.bz;
}

// fwd = SL(field, +dt) with the current velocity, for one field (NF = 1,
// field id f0) or density and temperature together (NF = 2).
template <typename T, int NF>
__global__ __launch_bounds__(kThreads) void kb_sl_forward(const BInst<T> *__restrict__ I,
                                                          const BItem *__restrict__ items,
                                                          double dt, Slots sl, int f0) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    int i, j, k;
    cell_of<T>(it, i, j, k);
    const int slots[5] = {sl.u, sl.v, sl.w, sl.rho, sl.t};
    const T *src0 = in.f[slots[f0]][f0];
    const T *src1 = NF == 2 ? in.f[sl.t][4] : nullptr;
    body_sl_forward<T, NF>(i, j, k, in.g, (T)(dt / in.h), in.f[sl.u][0], in.f[sl.v][1],
                           in.f[sl.w][2], src0, src1, in.fwd0, in.fwd1);
}

// MacCormack correction into the other slot of field f0 (and of
// temperature when NF = 2).
template <typename T, int NF>
__global__ __launch_bounds__(kThreads) void kb_mac_correct(const BInst<T> *__restrict__ I,
                                                           const BItem *__restrict__ items,
                                                           double dt, Slots sl, int f0,
                                                           double diss0, double diss1,
                                                           int clamp0) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    int i, j, k;
    cell_of<T>(it, i, j, k);
    const int slots[5] = {sl.u, sl.v, sl.w, sl.rho, sl.t};
    const int s0 = slots[f0];
    body_mac_correct<T, NF>(i, j, k, in.g, (T)(dt / in.h), in.f[sl.u][0], in.f[sl.v][1],
                            in.f[sl.w][2], in.f[s0][f0], in.fwd0, in.f[1 - s0][f0], (T)diss0,
                            NF == 2 ? in.f[sl.t][4] : nullptr, in.fwd1,
                            NF == 2 ? in.f[1 - sl.t][4] : nullptr, (T)diss1, clamp0);
}

// Sphere source: one workgroup per 256 cells of an instance's bounding box.
template <typename T>
__global__ __launch_bounds__(kThreads) void kb_sphere(const BInst<T> *__restrict__ I,
                                                      const int2 *__restrict__ items, Slots sl) {
    const int2 it = items[blockIdx.x];  // (instance, first box cell)
    const BInst<T> &in = I[it.x];
    const int a = it.y + tid_lin();
    const int ni = in.sn[0], nj = in.sn[1], nk = in.sn[2];
    if (a >= ni * nj * nk) return;
    const int i = in.slo[0] + a % ni, j = in.slo[1] + (a / ni) % nj, k = in.slo[2] + a / (ni * nj);
    const double dx = (double)i - in.sc[0], dy = (double)j - in.sc[1], dz = (double)k - in.sc[2];
    if (dx * dx + dy * dy + dz * dz < in.sr2) {
        const int id = (k * in.g.ny + j) * in.g.nx + i;
        in.f[sl.rho][3][id] = (T)in.srho;
        in.f[sl.t][4][id] = (T)in.stemp;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void kb_curl_mag(const BInst<T> *__restrict__ I,
                                                        const BItem *__restrict__ items,
                                                        Slots sl) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    int i, j, k;
    cell_of<T>(it, i, j, k);
    const SphereSrc none = {0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    body_curl_mag<T>(i, j, k, in.g, (T)(0.5 / in.h), in.f[sl.u][0], in.f[sl.v][1],
                     in.f[sl.w][2], in.curl, none, nullptr, nullptr);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void kb_confine_buoy_bc(const BInst<T> *__restrict__ I,
                                                               const BItem *__restrict__ items,
                                                               double dt, Slots sl) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    int i, j, k;
    cell_of<T>(it, i, j, k);
    body_confine_buoy_bc<T>(i, j, k, in.g, (T)dt, (T)(0.5 / in.h), (T)(in.eps * in.h),
                            (T)in.beta, (T)in.alpha, in.f[sl.u][0], in.f[sl.v][1], in.f[sl.w][2],
                            in.curl, in.f[sl.rho][3], in.f[sl.t][4], in.f[1 - sl.u][0],
                            in.f[1 - sl.v][1], in.f[1 - sl.w][2]);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void kb_gradient_bc(const BInst<T> *__restrict__ I,
                                                           const BItem *__restrict__ items,
                                                           Slots sl) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    int i, j, k;
    cell_of<T>(it, i, j, k);
    body_gradient_bc<T>(i, j, k, in.g, (T)(1.0 / in.h), in.p, in.f[sl.u][0], in.f[sl.v][1],
                        in.f[sl.w][2]);
}

// with_div 0: keep the stored b (the autotuner's synthetic right-hand side).
template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void kb_div_cginit(const BInst<T> *__restrict__ I,
                                                          const BItem *__restrict__ items,
                                                          Slots sl, int with_div) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    body_div_cginit<T, ZU>(col_at<ZU>(in.g, it.bx, it.by, it.bz), it.local, in.g, (T)in.h,
                           with_div,
                           in.f[sl.u][0], in.f[sl.v][1], in.f[sl.w][2], in.p, in.b, in.r, in.pb,
                           in.pc, in.st);
}

template <typename T, int ZU, bool FIRST>
__global__ __launch_bounds__(kThreads) void kb_cg_dir(const BInst<T> *__restrict__ I,
                                                      const BItem *__restrict__ items, int iter,
                                                      double rel_tol2, int *done_mirror,
                                                      int solve_id) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    body_cg_dir<T, ZU, FIRST>(col_at<ZU>(in.g, it.bx, it.by, it.bz), it.local, in.nb, in.g, iter,
                              rel_tol2, in.pb, in.pc, in.st, in.r, in.d[(iter + 1) & 1],
                              in.d[iter & 1], in.pa, done_mirror + it.inst, solve_id);
}

template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void kb_cg_update(const BInst<T> *__restrict__ I,
                                                         const BItem *__restrict__ items,
                                                         int iter, int *done_mirror,
                                                         int solve_id) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    body_cg_update<T, ZU>(col_at<ZU>(in.g, it.bx, it.by, it.bz), it.local, in.nb, in.g, iter,
                          in.pa, in.st, in.d[iter & 1], in.p, in.r, in.pb,
                          done_mirror + it.inst, solve_id);
}

// Four-wide CG launches over every instance's tiles (FP32, all nx % 4 == 0):
// the items list each instance's four-wide tiles in the order of the
// single-grid launch, so `local` is the same partial-sum slot and a batch
// computes bit for bit what the instances compute one at a time.
// PEND / PM: p is brought up to date on odd iterations only, as in the
// single-grid four-wide CG (same arithmetic, so batches stay bit-identical
// to the instances run one at a time).  VZ: z planes per thread of the
// tiling (the table's vt).
template <int VZ, bool FIRST, int NQ, bool PEND>
__global__ __launch_bounds__(kThreads) void kb_cg_dir4(const BInst<float> *__restrict__ I,
                                                       const BItem *__restrict__ items, int iter,
                                                       double rel_tol2, int *done_mirror,
                                                       int solve_id) {
    const BItem it = items[blockIdx.x];
    const BInst<float> &in = I[it.inst];
    if (!FIRST) {  // an instance that stopped returns before loading anything (the
                   // body's own test: stop word <= 2 iter, written by an earlier launch)
        const int done = in.st->done;
        if (done != 0 && done <= 2 * iter) return;
    }
    body_cg_dir4<VZ, FIRST, NQ, kStWT, PEND>(vcol_at<VZ>(in.g, in.vt, it.bx, it.by, it.bz),
                                            it.local, in.g, iter, rel_tol2,
                                            FIRST ? in.nb : in.nbv, in.pb, in.pc, in.st, in.r,
                                            in.d[(iter + 1) & 1], in.d[iter & 1], in.pa,
                                            done_mirror + it.inst, solve_id, in.p);
}

template <int VZ, int NQ, int PM>
__global__ __launch_bounds__(kThreads) void kb_cg_update4(const BInst<float> *__restrict__ I,
                                                          const BItem *__restrict__ items,
                                                          int iter, int *done_mirror,
                                                          int solve_id) {
    const BItem it = items[blockIdx.x];
    const BInst<float> &in = I[it.inst];
    {  // an instance that stopped returns before loading anything (stop word <= 2 iter + 1)
        const int done = in.st->done;
        if (done != 0 && done <= 2 * iter + 1) return;
    }
    body_cg_update4<VZ, NQ, kStWT, PM>(vcol_at<VZ>(in.g, in.vt, it.bx, it.by, it.bz), it.local,
                                      in.nbv, in.g, iter, in.pa, in.st, in.d[iter & 1], in.p,
                                      in.r, in.pb, done_mirror + it.inst, solve_id,
                                      in.d[(iter + 1) & 1]);
}

// delta_0 = r_0 . A r_0 partials and w_0 = A r_0 of every instance (as
// k_cg_rar on the initialisation tiling).
template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void kb_cg_rar(const BInst<T> *__restrict__ I,
                                                      const BItem *__restrict__ items) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    body_cg_rar<T, ZU>(col_at<ZU>(in.g, it.bx, it.by, it.bz), it.local, in.g, in.r, in.pe,
                       in.w[0]);
}

// One-launch CG iteration `iter` (cg_cgs4.hpp) over the items of a group of
// instances whose tilings share the launch's class RG and partial-sum width
// NQ: the same body, buffers and partial-sum slots as the single-grid
// launch, so each instance computes bit for bit what it computes alone.  A
// converged instance's tiles return before loading anything (its stop word
// holds the stopping launch + 1; the stopping launch itself never sees a
// value at or below its own index).
template <bool FIRST, int NQ, int PH, int RG>
__global__ __launch_bounds__(kThreads) void kb_cg_cgs4(const BInst<float> *__restrict__ I,
                                                       const BItem *__restrict__ items,
                                                       int iter, double rel_tol2,
                                                       int *done_mirror, int solve_id) {
    const BItem it = items[blockIdx.x];
    const BInst<float> &in = I[it.inst];
    if (!FIRST) {
        const int done = in.st->done;
        if (done != 0 && done <= iter) return;
    }
    VCol c;
    if constexpr (RG != 0)
        c = vcol_regular_at<4>(in.g, in.vt4, it.by, it.bz);
    else
        c = vcol_at<4>(in.g, in.vt4, it.bx, it.by, it.bz);
    CgRing ring;
#pragma unroll
    for (int q = 0; q < kCgRingMax; ++q) ring.r[q] = in.ring[q];
    const bool odd = iter & 1;
    body_cg_cgs4<4, FIRST, NQ, kStWT, kCgsCycle, PH, RG>(
        c, it.local, in.g, iter, rel_tol2, FIRST ? in.nb : in.nb4, odd ? in.pa : in.pb,
        odd ? in.pf : in.pe, in.pc, odd ? in.pb : in.pa, odd ? in.pe : in.pf, in.st, ring,
        in.w[iter & 1], in.w[(iter + 1) & 1], in.d[0], in.p, done_mirror + it.inst, solve_id);
}

template <int NQ, int RG, int... PH>
inline const void *kb_cgs4_phase(int ph, std::integer_sequence<int, PH...>) {
    const void *fns[] = {(const void *)kb_cg_cgs4<false, NQ, PH, RG>...};
    return fns[ph];
}

template <int NQ, int RG>
inline const void *kb_cgs4_pick(bool first, int ph) {
    if (first) return (const void *)kb_cg_cgs4<true, NQ, 0, RG>;
    return kb_cgs4_phase<NQ, RG>(ph, std::make_integer_sequence<int, kCgsCycle>{});
}

// The launch's class as the single grid chooses it (cgs4_kernel with
// write-through stores): nparts partial sums, tiling class rg.
inline int kb_cgs4_class(int nparts, int rg) {
    if (rg != 0 && nparts <= 4 * kThreads) return 10 * rg + (nparts <= 2 * kThreads ? 2 : 4);
    return nparts > 4 * kThreads ? 16 : 4;
}

inline const void *kb_cgs4_kernel(bool first, int ph, int cls) {
    switch (cls) {
    case 12: return kb_cgs4_pick<2, 1>(first, ph);
    case 14: return kb_cgs4_pick<4, 1>(first, ph);
    case 22: return kb_cgs4_pick<2, 2>(first, ph);
    case 24: return kb_cgs4_pick<4, 2>(first, ph);
    case 16: return kb_cgs4_pick<16, 0>(first, ph);
    default: return kb_cgs4_pick<4, 0>(first, ph);
    }
}

// Closes, per instance, a four-wide solve that an odd iteration cap ended
// after an even launch: p += alpha_last d_last over the instance's tiles.
template <int VZ>
__global__ __launch_bounds__(kThreads) void kb_cg_p_tail(const BInst<float> *__restrict__ I,
                                                         const BItem *__restrict__ items,
                                                         int last_it) {
    const BItem it = items[blockIdx.x];
    const BInst<float> &in = I[it.inst];
    if (in.st->done) return;
    const VCol c = vcol_at<VZ>(in.g, in.vt, it.bx, it.by, it.bz);
    const float ta = (float)in.st->cg_alpha[last_it & 1];
    float4 *p4 = reinterpret_cast<float4 *>(in.p);
    const float4 *d4 = reinterpret_cast<const float4 *>(in.d[last_it & 1]);
#pragma unroll
    for (int a = 0; a < VZ; ++a) {
        const int k = c.k0 + a;
        if (c.valid && k < in.g.nz) {
            const int id4 = (k * in.g.sxy + c.off) >> 2;
            p4[id4] = fma4(ta, d4[id4], p4[id4]);
        }
    }
}

// One workgroup per instance: closes solves that ran out of iterations and
// logs the iteration count of step `step` (of *step_dev when it is given: a
// step captured in a graph reads its index from the device).  form: the
// solve's launches, 0 one cell per thread, 1 four-wide (its r.r partials are
// nbv wide), 2 one-launch (nb4 wide, in pa after an odd count of launches).
template <typename T>
__global__ __launch_bounds__(kThreads) void kb_cg_finalize(const BInst<T> *__restrict__ I,
                                                           int max_iters, double rel_tol2,
                                                           int step, int form,
                                                           const int *__restrict__ step_dev) {
    const BInst<T> &in = I[blockIdx.x];
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    CgState *st = in.st;
    if (!block_flag(&st->done, flag_lds)) {
        const bool ran = max_iters > 0;
        const double *prr = (form == 2 && ran && (max_iters & 1)) ? in.pa : in.pb;
        const int nrr = !ran ? in.nb : form == 1 ? in.nbv : form == 2 ? in.nb4 : in.nb;
        const double rr = sum_partials(prr, nrr, lds);
        double bb = 0.0;
        if (max_iters == 0) bb = sum_partials(in.pc, in.nb, lds);
        if (tid_lin() == 0) {
            if (max_iters == 0) {
                st->b2 = bb;
                st->r2_init = rr;
            }
            st->iters = max_iters;
            st->r2_final = rr;
            st->converged = max_iters > 0 && (rr < 1e-20 ||
                                              (rel_tol2 > 0.0 && rr <= rel_tol2 * st->b2));
            st->done = 1;
        }
    }
    __syncthreads();
    if (tid_lin() == 0) in.iters[step_dev ? *step_dev : step] = st->iters;
}

template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void kb_telemetry_part(const BInst<T> *__restrict__ I,
                                                              const BItem *__restrict__ items,
                                                              Slots sl) {
    const BItem it = items[blockIdx.x];
    const BInst<T> &in = I[it.inst];
    const int kb = it.bz * ZU;
    body_telemetry_part<T>(it.bx * kBX + threadIdx.x, it.by * kBY + threadIdx.y, kb,
                           min(kb + ZU, in.g.nz), it.local, in.g, in.h, in.f[sl.u][0],
                           in.f[sl.v][1], in.f[sl.w][2], in.f[sl.rho][3], in.ptel);
}

// step_dev / times_dev (optional): the step index and the time axis on the
// device, for steps captured in a graph.
template <typename T>
__global__ __launch_bounds__(kThreads) void kb_telemetry_final(const BInst<T> *__restrict__ I,
                                                               int step, double time_sec,
                                                               const int *__restrict__ step_dev,
                                                               const double *__restrict__ times_dev) {
    const BInst<T> &in = I[blockIdx.x];
    const int s = step_dev ? *step_dev : step;
    body_telemetry_final(in.ptel, in.nb, step_dev ? times_dev[s] : time_sec,
                         in.tel + (size_t)s * 8);
}

// Last node of a captured step: advances the device step index.
__global__ void kb_step_advance(int *step_dev) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *step_dev += 1;
}

}  // namespace
}  // namespace aether
