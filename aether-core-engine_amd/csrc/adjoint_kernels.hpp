// adjoint_kernels.hpp -- reverse-mode kernels of the smoke step.
//
// Each kernel is the hand-written transpose of one forward stage of
// src/fluid.rs:598-611 linearised around a recorded forward state:
//
//   advection (src/fluid.rs:153-234): the MacCormack correction and the
//     forward semi-Lagrangian pass are reversed in turn.  The 8-cell clamp
//     acts as a selection mask (a clamped cell passes its adjoint to the
//     arg-min / arg-max corner), the near-wall branch and floor() cell
//     choice are piecewise constant, and the trilinear weights are
//     differentiated with respect to the back-traced position, which gives
//     the velocity adjoint.  Gathers become scatters, accumulated in 64-bit
//     fixed point (below), so the result does not depend on the order in
//     which threads add.
//   confinement + buoyancy (src/fluid.rs:256-330): three gather passes
//     (N x w and the normalisation of grad|w|, then the transpose of the
//     central differences of |w| and of the curl); deterministic.
//   projection (src/fluid.rs:496-529): self-adjoint (the wall-masked MAC
//     gradient is minus the transpose of the masked divergence and A is
//     symmetric), so the forward projection is applied to the adjoint
//     velocity by the host; no kernel here.
#pragma once
#include "fluid_kernels.hpp"

namespace aether {
namespace {

// Fixed-point scatter accumulation.  A contribution v is rounded to the
// nearest multiple of 2^-F and added as a 64-bit integer.  Integer addition
// is associative (wrap-around included), so a cell's sum is the same
// whatever order the threads add in: the adjoint is reproducible bit for
// bit.  F is set per scatter from a bound B on the sum of the absolute
// values of all contributions of that scatter (so on every cell's sum and
// every partial sum): F = 61 - ceil(log2 B), which keeps |sum| * 2^F below
// 2^61.  Each contribution is then exact to B * 2^-62 in absolute terms.
struct FixScale {
    double up;    // 2^F
    double down;  // 2^-F
};
typedef unsigned long long fixacc_t;

__device__ __forceinline__ long long fix_q(double v, double up) { return __double2ll_rn(v * up); }

__device__ __forceinline__ void fix_add(fixacc_t *p, double v, double up) {
    const long long q = fix_q(v, up);
    if (q) atomicAdd(p, (fixacc_t)q);
}

__device__ __forceinline__ double fix_value(fixacc_t a, double down) {
    return (double)(long long)a * down;
}

// Sum of |x| over n values (FP64 partials, one per workgroup).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_abs_sum(const T *__restrict__ x, long n,
                                                      double *__restrict__ part) {
    __shared__ double lds[8];
    double acc[1] = {0.0};
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads)
        acc[0] += fabs((double)x[a]);
    block_sum<1>(acc, lds);
    if (tid_lin() == 0) part[blockIdx.x] = acc[0];
}

// The scale of a scatter whose contributions sum in absolute value to at
// most factor * (sum of the nb partials).
__global__ __launch_bounds__(kThreads) void k_fix_scale(const double *__restrict__ part, int nb,
                                                        double factor, FixScale *out) {
    __shared__ double lds[8];
    const double s = sum_partials(part, nb, lds);
    if (tid_lin() == 0) {
        const double b = factor * s * (1.0 + 1e-6);
        int f = 0;
        if (b > 0.0 && b < 1e300) {
            int e = 0;
            frexp(b, &e);  // b < 2^e
            f = 61 - e;
        }
        f = max(-1000, min(1000, f));
        out->up = ldexp(1.0, f);
        out->down = ldexp(1.0, -f);
    }
}

// axis_loc plus the flag "position clamped" (zero derivative along the axis).
__device__ __forceinline__ void axis_loc_c(int i, double sv, int n, int &i0, double &t, bool &nw,
                                           bool &clamped) {
    const double x = (double)i - sv;
    clamped = x < 0.5 || x > (double)n - 1.500001;
    axis_loc(i, sv, n, i0, t, nw);
}
__device__ __forceinline__ void axis_loc_c(int i, float sv, int n, int &i0, float &t, bool &nw,
                                           bool &clamped) {
    const float o = -sv, fi = (float)i;
    clamped = (o < 0.5f - fi) || (o > (float)(n - 1 - i) - 0.500001f);
    axis_loc(i, sv, n, i0, t, nw);
}

template <typename T>
struct Loc {
    int i0, j0, k0;
    T tx, ty, tz;
    bool cx, cy, cz, nw;
};

template <typename T>
__device__ __forceinline__ Loc<T> locate(const Dims &g, int i, int j, int k, T su, T sv, T sw) {
    Loc<T> L;
    bool nwx, nwy, nwz;
    axis_loc_c(i, su, g.nx, L.i0, L.tx, nwx, L.cx);
    axis_loc_c(j, sv, g.ny, L.j0, L.ty, nwy, L.cy);
    axis_loc_c(k, sw, g.nz, L.k0, L.tz, nwz, L.cz);
    L.nw = nwx || nwy || nwz;
    return L;
}

// The 8 corner values and offsets of a location, corner order (di, dj, dk)
// with di fastest (the order of src/fluid.rs:140-148).
template <typename T>
__device__ __forceinline__ void corners(const Dims &g, const T *__restrict__ f, const Loc<T> &L,
                                        T (&c)[8], int (&ids)[8]) {
    const int b = (L.k0 * g.ny + L.j0) * g.nx + L.i0;
    const int o[8] = {0, 1, g.nx, g.nx + 1, g.sxy, g.sxy + 1, g.sxy + g.nx, g.sxy + g.nx + 1};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        ids[a] = b + o[a];
        c[a] = f[ids[a]];
    }
}

template <typename T>
__device__ __forceinline__ void weights(const Loc<T> &L, T (&w)[8]) {
    const T one = T(1);
    const T ax = one - L.tx, ay = one - L.ty, az = one - L.tz;
    w[0] = ax * ay * az;
    w[1] = L.tx * ay * az;
    w[2] = ax * L.ty * az;
    w[3] = L.tx * L.ty * az;
    w[4] = ax * ay * L.tz;
    w[5] = L.tx * ay * L.tz;
    w[6] = ax * L.ty * L.tz;
    w[7] = L.tx * L.ty * L.tz;
}

// d(trilinear)/d(position) along x, y, z; zero along a clamped axis.
template <typename T>
__device__ __forceinline__ void grad_pos(const Loc<T> &L, const T (&c)[8], T &gx, T &gy, T &gz) {
    const T one = T(1);
    const T ax = one - L.tx, ay = one - L.ty, az = one - L.tz;
    gx = ((c[1] - c[0]) * ay + (c[3] - c[2]) * L.ty) * az +
         ((c[5] - c[4]) * ay + (c[7] - c[6]) * L.ty) * L.tz;
    gy = ((c[2] - c[0]) * ax + (c[3] - c[1]) * L.tx) * az +
         ((c[6] - c[4]) * ax + (c[7] - c[5]) * L.tx) * L.tz;
    gz = (c[4] - c[0]) * ax * ay + (c[5] - c[1]) * L.tx * ay + (c[6] - c[2]) * ax * L.ty +
         (c[7] - c[3]) * L.tx * L.ty;
    if (L.cx) gx = T(0);
    if (L.cy) gy = T(0);
    if (L.cz) gz = T(0);
}

// Scatter targets of the reverse advection kernels.  GlobalAcc adds straight
// to a fixed-point array in global memory; ScatterWindow (below) first
// accumulates in LDS.  Both take the cell's index and coordinates.
struct GlobalAcc {
    fixacc_t *p;
    double up;
    __device__ __forceinline__ void add(int gid, int, int, int, double v) const {
        fix_add(p + gid, v, up);
    }
};

// Reverse of the MacCormack correction (src/fluid.rs:186-228) at one cell:
// lo_adj is the adjoint of the advected field there.  Scatters to the
// adjoints of the field before the step (asrc) and of the forward SL result
// (afwd), and returns the own-cell adjoint of the trace velocity in
// vu, vv, vw.
template <typename T, typename AS, typename AF>
__device__ __forceinline__ void adj_mac_cell(const Dims &g, int i, int j, int k, int id, T s,
                                             T diss, int clamp0, const T *__restrict__ src,
                                             const T *__restrict__ fwd, const T *__restrict__ U,
                                             const T *__restrict__ V, const T *__restrict__ W,
                                             T lo_adj, const AS &asrc, AF &afwd, T &vu, T &vv,
                                             T &vw) {
    vu = vv = vw = T(0);
    if (lo_adj == T(0)) return;
    const T su = s * U[id], sv = s * V[id], sw = s * W[id];
    const Loc<T> Lb = locate<T>(g, i, j, k, su, sv, sw);
    const T f = fwd[id];
    if (Lb.nw) {
        if (clamp0 && f * diss < T(0)) return;  // max(0) selected 0
        afwd.add(id, i, j, k, (double)(diss * lo_adj));
        return;
    }
    const Loc<T> Lf = locate<T>(g, i, j, k, -su, -sv, -sw);
    T cf[8], cs[8];
    int idf[8], ids[8];
    corners<T>(g, fwd, Lf, cf, idf);
    corners<T>(g, src, Lb, cs, ids);
    const T bwd = lerp3<T>(cf[0], cf[1], cf[2], cf[3], cf[4], cf[5], cf[6], cf[7], Lf.tx, Lf.ty,
                           Lf.tz);
    const T mc = f + T(0.5) * (src[id] - bwd);
    T lo = cs[0], hi = cs[0];
    int alo = 0, ahi = 0;
#pragma unroll
    for (int a = 1; a < 8; ++a) {
        if (cs[a] < lo) { lo = cs[a]; alo = a; }
        if (cs[a] > hi) { hi = cs[a]; ahi = a; }
    }
    T val = mc;
    if (val < lo) val = lo;
    if (val > hi) val = hi;
    if (clamp0 && val * diss < T(0)) return;
    const T gadj = diss * lo_adj;
    if (mc < lo) {
        asrc.add(ids[alo], Lb.i0 + (alo & 1), Lb.j0 + ((alo >> 1) & 1), Lb.k0 + (alo >> 2),
                 (double)gadj);
    } else if (mc > hi) {
        asrc.add(ids[ahi], Lb.i0 + (ahi & 1), Lb.j0 + ((ahi >> 1) & 1), Lb.k0 + (ahi >> 2),
                 (double)gadj);
    } else {
        afwd.add(id, i, j, k, (double)gadj);
        asrc.add(id, i, j, k, (double)(T(0.5) * gadj));
        T w[8];
        weights<T>(Lf, w);
        const T hb = T(-0.5) * gadj;
#pragma unroll
        for (int a = 0; a < 8; ++a)
            afwd.add(idf[a], Lf.i0 + (a & 1), Lf.j0 + ((a >> 1) & 1), Lf.k0 + (a >> 2),
                     (double)(hb * w[a]));
        // bwd position = cell + s * velocity
        T gx, gy, gz;
        grad_pos<T>(Lf, cf, gx, gy, gz);
        vu = hb * gx * s;
        vv = hb * gy * s;
        vw = hb * gz * s;
    }
}

// Reverse of the forward semi-Lagrangian pass fwd = SL(src, +dt)
// (src/fluid.rs:153-167) at one cell: lf is the adjoint of fwd there.
// Scatters to the 8 back-trace corners of src and adds the own-cell
// velocity adjoint (position = cell - s * velocity) to vu, vv, vw.
template <typename T, typename AS>
__device__ __forceinline__ void adj_sl_cell(const Dims &g, int i, int j, int k, int id, T s,
                                            const T *__restrict__ src, const T *__restrict__ U,
                                            const T *__restrict__ V, const T *__restrict__ W,
                                            T lf, AS &asrc, T &vu, T &vv, T &vw) {
    if (lf == T(0)) return;
    const Loc<T> L = locate<T>(g, i, j, k, s * U[id], s * V[id], s * W[id]);
    T c[8], w[8];
    int ids[8];
    corners<T>(g, src, L, c, ids);
    weights<T>(L, w);
#pragma unroll
    for (int a = 0; a < 8; ++a)
        asrc.add(ids[a], L.i0 + (a & 1), L.j0 + ((a >> 1) & 1), L.k0 + (a >> 2),
                 (double)(lf * w[a]));
    T gx, gy, gz;
    grad_pos<T>(L, c, gx, gy, gz);
    vu += -lf * gx * s;
    vv += -lf * gy * s;
    vw += -lf * gz * s;
}

// One thread per cell.  acc_src / acc_fwd: fixed-point adjoints of the
// field before the step and of the forward SL result (zeroed by the
// caller); tU/tV/tW: the own-cell velocity adjoint, written at every cell.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_mac_correct(
    Dims g, T s, T diss, int clamp0, const T *__restrict__ src, const T *__restrict__ fwd,
    const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
    const T *__restrict__ lam_out, const FixScale *sc, fixacc_t *acc_src, fixacc_t *acc_fwd,
    T *__restrict__ tU, T *__restrict__ tV, T *__restrict__ tW) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    T vu = T(0), vv = T(0), vw = T(0);
    if (is_interior(g, i, j, k)) {
        const GlobalAcc as{acc_src, sc->up};
        GlobalAcc af{acc_fwd, sc->up};
        adj_mac_cell<T>(g, i, j, k, id, s, diss, clamp0, src, fwd, U, V, W, lam_out[id], as, af,
                        vu, vv, vw);
    }
    tU[id] = vu;
    tV[id] = vv;
    tW[id] = vw;
}

// One thread per cell: reads the forward adjoint from acc_fwd, scatters to
// acc_src and adds to the own-cell velocity adjoint tU/tV/tW.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_sl_forward(
    Dims g, T s, const T *__restrict__ src, const T *__restrict__ U, const T *__restrict__ V,
    const T *__restrict__ W, const FixScale *sc, const fixacc_t *__restrict__ acc_fwd,
    fixacc_t *acc_src, T *__restrict__ tU, T *__restrict__ tV, T *__restrict__ tW) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny || !is_interior(g, i, j, k)) return;
    const int id = (k * g.ny + j) * g.nx + i;
    const T lf = (T)fix_value(acc_fwd[id], sc->down);
    if (lf == T(0)) return;
    GlobalAcc as{acc_src, sc->up};
    T vu = tU[id], vv = tV[id], vw = tW[id];
    adj_sl_cell<T>(g, i, j, k, id, s, src, U, V, W, lf, as, vu, vv, vw);
    tU[id] = vu;
    tV[id] = vv;
    tW[id] = vw;
}

// ---------------------------------------------------------------------------
// Windowed scatters.  A workgroup owns a WTX x WTY x WTZ tile of cells; the
// trilinear corners its cells scatter to are accumulated in an LDS window
// covering the tile grown by H cells (H + 1 on the high side, for the second
// corner), then flushed to global memory with one atomic add per nonzero
// window cell, consecutive in x.  A corner outside the window (a trace longer
// than H cells) goes straight to global memory.  With traces of up to H
// cells this replaces 8 scattered global atomics per cell by ~4 row-coalesced
// ones (window / tile = 39*15*15 / 2048 for H = 3).  Window and global
// memory hold the same fixed-point integers, so where a contribution is
// summed first does not change the result.
// ---------------------------------------------------------------------------
constexpr int WTX = 32, WTY = 8, WTZ = 8, WH = 3;
constexpr int WWX = WTX + 2 * WH + 1, WWY = WTY + 2 * WH + 1, WWZ = WTZ + 2 * WH + 1;
constexpr int WN = WWX * WWY * WWZ;
static_assert(WTX * WTY == kThreads, "one thread per column of the tile");

struct ScatterWindow {
    fixacc_t *lds;
    fixacc_t *global;
    double up;
    int x0, y0, z0;  // global cell of window entry 0
    __device__ __forceinline__ void init(fixacc_t *mem, fixacc_t *glob, double scale, int tx0,
                                         int ty0, int tz0) {
        lds = mem;
        global = glob;
        up = scale;
        x0 = tx0 - WH;
        y0 = ty0 - WH;
        z0 = tz0 - WH;
        for (int e = threadIdx.y * WTX + threadIdx.x; e < WN; e += kThreads) lds[e] = 0;
    }
    // cell (i, j, k) = global index gid gets v
    __device__ __forceinline__ void add(int gid, int i, int j, int k, double v) {
        const long long q = fix_q(v, up);
        if (!q) return;
        const unsigned a = (unsigned)(i - x0), b = (unsigned)(j - y0), c = (unsigned)(k - z0);
        if (a < (unsigned)WWX && b < (unsigned)WWY && c < (unsigned)WWZ)
            atomicAdd(lds + ((int)c * WWY + (int)b) * WWX + (int)a, (fixacc_t)q);
        else
            atomicAdd(global + gid, (fixacc_t)q);
    }
    __device__ __forceinline__ void flush(const Dims &g) const {
        for (int e = threadIdx.y * WTX + threadIdx.x; e < WN; e += kThreads) {
            const fixacc_t v = lds[e];
            if (v == 0) continue;
            const int a = e % WWX, rest = e / WWX;
            const int i = x0 + a, j = y0 + rest % WWY, k = z0 + rest / WWY;
            if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) continue;
            atomicAdd(global + (k * g.ny + j) * g.nx + i, v);
        }
    }
};

// k_adj_mac_correct with acc_fwd accumulated in a window (grid: tiles of
// WTX x WTY x WTZ cells, block WTX x WTY, each thread a z column).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_mac_correct_w(
    Dims g, T s, T diss, int clamp0, const T *__restrict__ src, const T *__restrict__ fwd,
    const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
    const T *__restrict__ lam_out, const FixScale *sc, fixacc_t *acc_src, fixacc_t *acc_fwd,
    T *__restrict__ tU, T *__restrict__ tV, T *__restrict__ tW) {
    __shared__ fixacc_t win_mem[WN];
    ScatterWindow win;
    const double up = sc->up;
    const int tx0 = blockIdx.x * WTX, ty0 = blockIdx.y * WTY, tz0 = blockIdx.z * WTZ;
    win.init(win_mem, acc_fwd, up, tx0, ty0, tz0);
    __syncthreads();
    const GlobalAcc as{acc_src, up};
    const int i = tx0 + threadIdx.x, j = ty0 + threadIdx.y;
    for (int kk = 0; kk < WTZ; ++kk) {
        const int k = tz0 + kk;
        if (i >= g.nx || j >= g.ny || k >= g.nz) continue;
        const int id = (k * g.ny + j) * g.nx + i;
        T vu = T(0), vv = T(0), vw = T(0);
        if (is_interior(g, i, j, k))
            adj_mac_cell<T>(g, i, j, k, id, s, diss, clamp0, src, fwd, U, V, W, lam_out[id], as,
                            win, vu, vv, vw);
        tU[id] = vu;
        tV[id] = vv;
        tW[id] = vw;
    }
    __syncthreads();
    win.flush(g);
}

// k_adj_sl_forward with acc_src accumulated in a window (same tiling).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_sl_forward_w(
    Dims g, T s, const T *__restrict__ src, const T *__restrict__ U, const T *__restrict__ V,
    const T *__restrict__ W, const FixScale *sc, const fixacc_t *__restrict__ acc_fwd,
    fixacc_t *acc_src, T *__restrict__ tU, T *__restrict__ tV, T *__restrict__ tW) {
    __shared__ fixacc_t win_mem[WN];
    ScatterWindow win;
    const double down = sc->down;
    const int tx0 = blockIdx.x * WTX, ty0 = blockIdx.y * WTY, tz0 = blockIdx.z * WTZ;
    win.init(win_mem, acc_src, sc->up, tx0, ty0, tz0);
    __syncthreads();
    const int i = tx0 + threadIdx.x, j = ty0 + threadIdx.y;
    for (int kk = 0; kk < WTZ; ++kk) {
        const int k = tz0 + kk;
        if (i >= g.nx || j >= g.ny || k >= g.nz || !is_interior(g, i, j, k)) continue;
        const int id = (k * g.ny + j) * g.nx + i;
        const T lf = (T)fix_value(acc_fwd[id], down);
        if (lf == T(0)) continue;
        T vu = tU[id], vv = tV[id], vw = tW[id];
        adj_sl_cell<T>(g, i, j, k, id, s, src, U, V, W, lf, win, vu, vv, vw);
        tU[id] = vu;
        tV[id] = vv;
        tW[id] = vw;
    }
    __syncthreads();
    win.flush(g);
}

// The reverse advection's results into the adjoint arrays, in a fixed
// order per cell: lam_src += acc_src, then lam_U += tU, lam_V += tV,
// lam_W += tW.  The arrays may alias (the w advection's source and trace
// velocity are the same field), so nothing here is __restrict__.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_flush(long n, const FixScale *sc,
                                                        const fixacc_t *acc_src, const T *tU,
                                                        const T *tV, const T *tW, T *lam_src,
                                                        T *lam_U, T *lam_V, T *lam_W) {
    const double down = sc->down;
    for (long q = (long)blockIdx.x * kThreads + tid_lin(); q < n; q += (long)gridDim.x * kThreads) {
        const T a = (T)fix_value(acc_src[q], down), bu = tU[q], bv = tV[q], bw = tW[q];
        lam_src[q] = lam_src[q] + a;
        lam_U[q] = lam_U[q] + bu;
        lam_V[q] = lam_V[q] + bv;
        lam_W[q] = lam_W[q] + bw;
    }
}

// Confinement/buoyancy reverse, pass 1 (src/fluid.rs:284-329): wall masks,
// identity part, buoyancy adjoints (density, temperature, beta), and for
// cells [2, n-3] the adjoints of the curl at the cell (lw) and of
// grad|w| (lg).  lam_new is the adjoint of the step's output velocity.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_confine1(
    Dims g, T dt, T inv_2h, T eps_h, T beta, T alpha, const T *__restrict__ u,
    const T *__restrict__ v, const T *__restrict__ w, const T *__restrict__ cmag,
    const T *__restrict__ temp, const T *__restrict__ lu_new, const T *__restrict__ lv_new,
    const T *__restrict__ lw_new, T *__restrict__ lu, T *__restrict__ lv, T *__restrict__ lw,
    T *__restrict__ lwx, T *__restrict__ lwy, T *__restrict__ lwz, T *__restrict__ lgx,
    T *__restrict__ lgy, T *__restrict__ lgz, T *__restrict__ lrho, T *__restrict__ ltemp,
    double *__restrict__ part_beta) {
    __shared__ double lds[8];
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    double acc[1] = {0.0};
    if (i < g.nx && j < g.ny) {
        const int id = (k * g.ny + j) * g.nx + i;
        const T mu = bc_u(g, i, j, k) ? T(0) : lu_new[id];
        const T mv = bc_v(g, i, j, k) ? T(0) : lv_new[id];
        const T mw = bc_w(g, i, j, k) ? T(0) : lw_new[id];
        lu[id] += mu;
        lv[id] += mv;
        lw[id] += mw;
        if (is_interior(g, i, j, k)) {
            // lift = beta*T - alpha*rho, v += dt*lift
            lrho[id] += -alpha * dt * mv;
            ltemp[id] += beta * dt * mv;
            acc[0] = (double)dt * (double)temp[id] * (double)mv;
        }
        T ax = 0, ay = 0, az = 0, gxo = 0, gyo = 0, gzo = 0;
        if (i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k >= 2 && k <= g.nz - 3) {
            const T fxa = dt * mu, fya = dt * mv, fza = dt * mw;  // adjoint of the force
            const T gx = (cmag[id + 1] - cmag[id - 1]) * inv_2h;
            const T gy = (cmag[id + g.nx] - cmag[id - g.nx]) * inv_2h;
            const T gz = (cmag[id + g.sxy] - cmag[id - g.sxy]) * inv_2h;
            const T gl = sqrt(gx * gx + gy * gy + gz * gz);
            const T glen = gl + T(1e-10);
            const T nx = gx / glen, ny = gy / glen, nz = gz / glen;
            T ox, oy, oz;
            curl_at<T>(g, u, v, w, id, inv_2h, ox, oy, oz);
            // f = eps_h * (N x w):  df.a/dw = eps_h * (a x N),  df.a/dN = eps_h * (w x a)
            ax = eps_h * (fya * nz - fza * ny);
            ay = eps_h * (fza * nx - fxa * nz);
            az = eps_h * (fxa * ny - fya * nx);
            const T lnx = eps_h * (oy * fza - oz * fya);
            const T lny = eps_h * (oz * fxa - ox * fza);
            const T lnz = eps_h * (ox * fya - oy * fxa);
            // N = g / (|g| + 1e-10)
            gxo = lnx / glen;
            gyo = lny / glen;
            gzo = lnz / glen;
            if (gl > T(0)) {
                const T dot = (gx * lnx + gy * lny + gz * lnz) / (gl * glen * glen);
                gxo -= gx * dot;
                gyo -= gy * dot;
                gzo -= gz * dot;
            }
        }
        lwx[id] = ax;
        lwy[id] = ay;
        lwz[id] = az;
        lgx[id] = gxo;
        lgy[id] = gyo;
        lgz[id] = gzo;
    }
    block_sum<1>(acc, lds);
    if (tid_lin() == 0)
        part_beta[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = acc[0];
}

// Pass 2: adjoint of |w| from grad|w| (central differences), then the full
// adjoint of w at interior cells: lw += lam_|w| * w / |w|.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_confine2(
    Dims g, T inv_2h, const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ cmag, const T *__restrict__ lgx, const T *__restrict__ lgy,
    const T *__restrict__ lgz, T *__restrict__ lwx, T *__restrict__ lwy, T *__restrict__ lwz) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny || !is_interior(g, i, j, k)) return;
    const int id = (k * g.ny + j) * g.nx + i;
    // g_a(x) = (|w|[x + e_a] - |w|[x - e_a]) * inv_2h, defined on [2, n-3]
    const T lm = inv_2h * ((lgx[id - 1] - lgx[id + 1]) + (lgy[id - g.nx] - lgy[id + g.nx]) +
                           (lgz[id - g.sxy] - lgz[id + g.sxy]));
    const T m = cmag[id];
    if (m > T(0)) {
        T ox, oy, oz;
        curl_at<T>(g, u, v, w, id, inv_2h, ox, oy, oz);
        lwx[id] += lm * ox / m;
        lwy[id] += lm * oy / m;
        lwz[id] += lm * oz / m;
    }
}

// Pass 3: transpose of the central-difference curl (src/fluid.rs:265-274):
// lam_u, lam_v, lam_w += curl^T lam_w (lam_w is zero off the interior).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_confine3(
    Dims g, T inv_2h, const T *__restrict__ lwx, const T *__restrict__ lwy,
    const T *__restrict__ lwz, T *__restrict__ lu, T *__restrict__ lv, T *__restrict__ lw) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    auto at = [&](const T *a, int di, int dj, int dk) -> T {
        const int ii = i + di, jj = j + dj, kk = k + dk;
        if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) return T(0);
        return a[(kk * g.ny + jj) * g.nx + ii];
    };
    // ox = (w[y+1]-w[y-1]) h2 - (v[z+1]-v[z-1]) h2
    // oy = (u[z+1]-u[z-1]) h2 - (w[x+1]-w[x-1]) h2
    // oz = (v[x+1]-v[x-1]) h2 - (u[y+1]-u[y-1]) h2
    lu[id] += inv_2h * ((at(lwy, 0, 0, -1) - at(lwy, 0, 0, 1)) -
                        (at(lwz, 0, -1, 0) - at(lwz, 0, 1, 0)));
    lv[id] += inv_2h * ((at(lwz, -1, 0, 0) - at(lwz, 1, 0, 0)) -
                        (at(lwx, 0, 0, -1) - at(lwx, 0, 0, 1)));
    lw[id] += inv_2h * ((at(lwx, 0, -1, 0) - at(lwx, 0, 1, 0)) -
                        (at(lwy, -1, 0, 0) - at(lwy, 1, 0, 0)));
}

// Inflow patch: forward sets density and temperature to the control value
// in a box; the adjoint moves the box's adjoints into the control gradient
// and clears them (the overwritten values do not reach the loss).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_patch_set(Dims g, int3 lo, int3 n,
                                                        const double *__restrict__ ctrl,
                                                        T *__restrict__ rho,
                                                        T *__restrict__ temp) {
    const int a = blockIdx.x * kThreads + tid_lin();
    if (a >= n.x * n.y * n.z) return;
    const int i = lo.x + a % n.x, j = lo.y + (a / n.x) % n.y, k = lo.z + a / (n.x * n.y);
    const int id = (k * g.ny + j) * g.nx + i;
    rho[id] = (T)ctrl[a];
    temp[id] = (T)ctrl[a];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_patch_adj(Dims g, int3 lo, int3 n,
                                                        T *__restrict__ lrho, T *__restrict__ ltemp,
                                                        double *__restrict__ grad) {
    const int a = blockIdx.x * kThreads + tid_lin();
    if (a >= n.x * n.y * n.z) return;
    const int i = lo.x + a % n.x, j = lo.y + (a / n.x) % n.y, k = lo.z + a / (n.x * n.y);
    const int id = (k * g.ny + j) * g.nx + i;
    grad[a] += (double)lrho[id] + (double)ltemp[id];
    lrho[id] = T(0);
    ltemp[id] = T(0);
}

// Loss sum (rho - target)^2 (FP64 partials) and its adjoint 2 (rho - target).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_loss(const T *__restrict__ rho,
                                                   const double *__restrict__ target, long n,
                                                   T *__restrict__ lrho,
                                                   double *__restrict__ part) {
    __shared__ double lds[8];
    double acc[1] = {0.0};
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads) {
        const double d = (double)rho[a] - target[a];
        acc[0] += d * d;
        if (lrho) lrho[a] = (T)(2.0 * d);
    }
    block_sum<1>(acc, lds);
    if (tid_lin() == 0) part[blockIdx.x] = acc[0];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_copy(const T *__restrict__ a, long n,
                                                   T *__restrict__ b) {
    for (long q = (long)blockIdx.x * kThreads + tid_lin(); q < n; q += (long)gridDim.x * kThreads)
        b[q] = a[q];
}

}  // namespace
}  // namespace aether
