// fluid_kernels.hpp -- CDNA4 (gfx950) kernels of the smoke step.
//
// Every kernel is a template on the storage type T (float or double).  The
// double instantiation evaluates each expression in the order of the
// reference (src/fluid.rs), and the library is compiled with
// -ffp-contract=off, so that FP64 results agree with the reference to the
// last bit wherever no reduction is involved.  The float instantiation keeps
// the same algorithm but uses forms that lose less precision in FP32: the
// back-trace is carried as (cell index, FP32 offset) instead of an absolute
// coordinate, and the Laplacian is summed as differences (x_c - x_nb).
// Reductions (CG dot products, telemetry) always accumulate in FP64, in a
// fixed tree order, so every result is run-to-run deterministic.
//
// Layout: one array of nx*ny*nz values per field, idx = (k*ny + j)*nx + i
// (src/fluid.rs:88-90), x fastest.  u/v/w hold the reference's dual-use
// velocity: cell-centred for advection/curl, faces at i-1/2 for the MAC
// divergence and gradient.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

namespace aether {
namespace {  // kernels and device helpers: private to each translation unit

constexpr int kBX = 64;  // one wavefront spans 64 consecutive x cells
constexpr int kBY = 4;   // four wavefronts per workgroup
constexpr int kThreads = kBX * kBY;

// Tiling of the z-marching reduction launches (divergence, telemetry, max
// residual): a workgroup covers a kBX x kBY column and marches zchunk planes,
// with the chunk chosen so that a grid gives ~1024 partial sums (more when the
// columns alone exceed that).  Partial slot = (z chunk * gy + y) * gx + x.
struct RedTiling {
    int gx, gy, nch, zchunk, nb;
};
inline RedTiling red_tiling(int nx, int ny, int nz) {
    RedTiling t;
    t.gx = (nx + kBX - 1) / kBX;
    t.gy = (ny + kBY - 1) / kBY;
    const int cols = t.gx * t.gy;
    int nch = std::max(1, std::min(nz, (1024 + cols - 1) / cols));
    t.zchunk = (nz + nch - 1) / nch;
    t.nch = (nz + t.zchunk - 1) / t.zchunk;
    t.nb = cols * t.nch;
    return t;
}

struct Dims {
    int nx, ny, nz;
    int sxy;  // nx*ny
};

// Device state of one pressure solve (written only by workgroup 0 of a
// kernel, read by later kernels on the same stream).
// Longest update cycle of the one-launch four-wide CG: its residual ring
// holds up to this many arrays, and CgState keeps alpha and beta of this
// many iterations.
constexpr int kCgRingMax = 8;

struct CgState {
    double delta[2];  // r.r of the current iteration, double-buffered by iteration parity
    double b2;        // ||b||^2
    double r2_init;   // ||b - A p0||^2
    double r2_final;  // recursive residual at exit
    int done;  // nonzero once the solve has stopped (the one-launch four-wide
               // form stores its stopping launch + 1)
    int iters;
    int converged;
    int pad;
    double cg_gamma[2];           // CG scalars by iteration: r.r (two slots, by parity),
    double cg_alpha[kCgRingMax];  // alpha and beta (by it mod kCgRingMax: the
    double cg_beta[kCgRingMax];   // one-launch CG reads up to K - 1 iterations back)
};

__device__ __forceinline__ int tid_lin() { return threadIdx.y * kBX + threadIdx.x; }

// Lane exchange inside 16-lane rows through DPP (no LDS round trip):
// kXor1 / kXor2 are quad_perm [1,0,3,2] / [2,3,0,1]; kMirror8 / kMirror16
// reverse each half-row / row.  Once the lanes of every 4-group (8-group)
// hold equal values, the 8- (16-) lane mirror pairs the same groups as an
// xor by 4 (8) would.
constexpr int kXor1 = 0xB1, kXor2 = 0x4E, kMirror8 = 0x141, kMirror16 = 0x140;

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double lane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Butterfly over the 64 lanes: pairs by xor 1, 2, 4, 8 inside each row
// (DPP), then rows (0 + 1) + (2 + 3).  Every lane ends with the
// bit-identical total.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<kXor1>(v);
    v += dpp_f64<kXor2>(v);
    v += dpp_f64<kMirror8>(v);
    v += dpp_f64<kMirror16>(v);
    return (lane_f64(v, 0) + lane_f64(v, 16)) + (lane_f64(v, 32) + lane_f64(v, 48));
}

__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_f64<kXor1>(v));
    v = fmax(v, dpp_f64<kXor2>(v));
    v = fmax(v, dpp_f64<kMirror8>(v));
    v = fmax(v, dpp_f64<kMirror16>(v));
    return fmax(fmax(lane_f64(v, 0), lane_f64(v, 16)), fmax(lane_f64(v, 32), lane_f64(v, 48)));
}

// Sum NV values over the workgroup; every thread receives the totals.
// lds must hold 4*NV doubles.
// REUSE: lds may be written again later in the kernel, so a second barrier
// keeps a fast wave from overwriting it while a slow one still reads it.
// Callers that give every reduction its own lds slots pass false.
template <int NV, bool REUSE = true>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds) {
    const int t = tid_lin(), lane = t & 63, wv = t >> 6;
#pragma unroll
    for (int a = 0; a < NV; ++a) v[a] = wave_sum(v[a]);
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < NV; ++a) lds[a * 4 + wv] = v[a];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NV; ++a)
        v[a] = ((lds[a * 4 + 0] + lds[a * 4 + 1]) + lds[a * 4 + 2]) + lds[a * 4 + 3];
    if (REUSE) __syncthreads();
}

__device__ __forceinline__ double block_max(double v, double *lds) {
    const int t = tid_lin(), lane = t & 63, wv = t >> 6;
    v = wave_max(v);
    if (lane == 0) lds[wv] = v;
    __syncthreads();
    v = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
    __syncthreads();
    return v;
}

// Reads a device flag once per workgroup so that every wave takes the same
// branch even if another workgroup updates the flag meanwhile.
__device__ __forceinline__ bool block_flag(const int *flag, int *lds_i) {
    if (tid_lin() == 0) lds_i[0] = __atomic_load_n(flag, __ATOMIC_RELAXED);
    __syncthreads();
    const bool f = lds_i[0] != 0;
    return f;
}

// Sum of nb per-workgroup partials, in the same order in every workgroup.
__device__ __forceinline__ double sum_partials(const double *part, int nb, double *lds) {
    double acc[1] = {0.0};
    for (int a = tid_lin(); a < nb; a += kThreads) acc[0] += part[a];
    block_sum<1>(acc, lds);
    return acc[0];
}

// Partial sums of the previous kernel, loaded early (before the data loads of
// the caller) so that both memory round trips overlap.  Up to NQ per thread
// are held in registers; any excess is summed in a loop afterwards.
template <int NQ>
struct EarlyPartialsT {
    double v[NQ];
    __device__ __forceinline__ void load(const double *__restrict__ part, int nb) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int a = tid_lin() + q * kThreads;
            const double x = part[min(a, nb - 1)];  // always a valid address
            v[q] = a < nb ? x : 0.0;
        }
    }
    // this thread's share of the total, in a fixed order
    __device__ __forceinline__ double local(const double *__restrict__ part, int nb) const {
        double acc;
        if constexpr (NQ == 8)
            acc = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
        else if constexpr (NQ == 4)
            acc = (v[0] + v[1]) + (v[2] + v[3]);
        else {
            acc = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc += v[q];
        }
        for (int a = tid_lin() + NQ * kThreads; a < nb; a += kThreads) acc += part[a];
        return acc;
    }
    // fixed-order total, identical in every workgroup
    __device__ __forceinline__ double total(const double *__restrict__ part, int nb,
                                            double *lds) const {
        double acc[1] = {local(part, nb)};
        block_sum<1>(acc, lds);
        return acc[0];
    }
};
using EarlyPartials = EarlyPartialsT<8>;

__device__ __forceinline__ bool is_interior(const Dims &g, int i, int j, int k) {
    return i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2 && k >= 1 && k <= g.nz - 2;
}

// Wall masks of src/fluid.rs:241-249.
__device__ __forceinline__ bool bc_u(const Dims &g, int i, int j, int k) {
    return i <= 1 || i + 1 >= g.nx || j == 0 || j + 1 >= g.ny || k == 0 || k + 1 >= g.nz;
}
__device__ __forceinline__ bool bc_v(const Dims &g, int i, int j, int k) {
    return j <= 1 || j + 1 >= g.ny || i == 0 || i + 1 >= g.nx || k == 0 || k + 1 >= g.nz;
}
__device__ __forceinline__ bool bc_w(const Dims &g, int i, int j, int k) {
    return k <= 1 || k + 1 >= g.nz || i == 0 || i + 1 >= g.nx || j == 0 || j + 1 >= g.ny;
}

// ---------------------------------------------------------------------------
// Back-trace location along one axis: position x = i - sv (sv = signed
// dt/h * velocity), clamped to [0.5, n - 1.500001] (src/fluid.rs:95-108),
// plus the near-wall test of src/fluid.rs:196-201.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void axis_loc(int i, double sv, int n, int &i0, double &t, bool &nw) {
    double x = (double)i - sv;
    nw = x < 1.5 || x > (double)n - 2.5;
    double xc = x;
    if (xc < 0.5) xc = 0.5;
    if (xc > (double)n - 1.500001) xc = (double)n - 1.500001;
    double f = floor(xc);
    i0 = (int)f;
    t = xc - f;
    // a NaN velocity passes both clamps; keep the gather inside the grid
    i0 = min(max(i0, 0), n - 2);
}

__device__ __forceinline__ void axis_loc(int i, float sv, int n, int &i0, float &t, bool &nw) {
    // x = i + o with o = -sv kept separately: comparisons against the clamp
    // and near-wall bounds are exact rewrites of the reference's tests.
    const float o = -sv;
    const float fi = (float)i;
    nw = (o < 1.5f - fi) || (o > ((float)n - 2.5f) - fi);
    if (o < 0.5f - fi) {
        i0 = 0;
        t = 0.5f;
    } else if (o > (float)(n - 1 - i) - 0.500001f) {
        i0 = n - 2;
        t = 0.499999f;
    } else {
        const float f = floorf(o);
        i0 = i + (int)f;
        t = o - f;
    }
    i0 = min(max(i0, 0), n - 2);  // NaN guard, as above
}

template <typename T>
__device__ __forceinline__ T lerp3(T c000, T c100, T c010, T c110, T c001, T c101, T c011,
                                   T c111, T tx, T ty, T tz) {
    // src/fluid.rs:119-127, same association
    const T one = T(1);
    T c00 = c000 * (one - tx) + c100 * tx;
    T c10 = c010 * (one - tx) + c110 * tx;
    T c01 = c001 * (one - tx) + c101 * tx;
    T c11 = c011 * (one - tx) + c111 * tx;
    T c0 = c00 * (one - ty) + c10 * ty;
    T c1 = c01 * (one - ty) + c11 * ty;
    return c0 * (one - tz) + c1 * tz;
}

template <typename T>
__device__ __forceinline__ T gather_lerp(const Dims &g, const T *__restrict__ f, int i0, int j0,
                                         int k0, T tx, T ty, T tz) {
    const int b = (k0 * g.ny + j0) * g.nx + i0;
    return lerp3<T>(f[b], f[b + 1], f[b + g.nx], f[b + g.nx + 1], f[b + g.sxy],
                    f[b + g.sxy + 1], f[b + g.sxy + g.nx], f[b + g.sxy + g.nx + 1], tx, ty, tz);
}

template <typename T>
__device__ __forceinline__ void gather_minmax(const Dims &g, const T *__restrict__ f, int i0,
                                              int j0, int k0, T &lo, T &hi) {
    const int b = (k0 * g.ny + j0) * g.nx + i0;
    T c[8] = {f[b], f[b + 1], f[b + g.nx], f[b + g.nx + 1], f[b + g.sxy], f[b + g.sxy + 1],
              f[b + g.sxy + g.nx], f[b + g.sxy + g.nx + 1]};
    lo = c[0];
    hi = c[0];
#pragma unroll
    for (int a = 1; a < 8; ++a) {
        lo = fmin(lo, c[a]);
        hi = fmax(hi, c[a]);
    }
}

// ---------------------------------------------------------------------------
// Advection (src/fluid.rs:153-234)
// ---------------------------------------------------------------------------

// Forward semi-Lagrangian pass of NF fields sharing one velocity:
// fwd = SL(src, +dt) on interior cells, 0 on the boundary layer.
template <typename T, int NF>
__device__ __forceinline__ void body_sl_forward(int i, int j, int k, Dims g, T s, const T *__restrict__ u,
                                                         const T *__restrict__ v,
                                                         const T *__restrict__ w,
                                                         const T *__restrict__ src0,
                                                         const T *__restrict__ src1,
                                                         T *__restrict__ fwd0,
                                                         T *__restrict__ fwd1) {
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    if (!is_interior(g, i, j, k)) {
        fwd0[id] = T(0);
        if (NF == 2) fwd1[id] = T(0);
        return;
    }
    int i0, j0, k0;
    T tx, ty, tz;
    bool nw;
    axis_loc(i, s * u[id], g.nx, i0, tx, nw);
    axis_loc(j, s * v[id], g.ny, j0, ty, nw);
    axis_loc(k, s * w[id], g.nz, k0, tz, nw);
    fwd0[id] = gather_lerp<T>(g, src0, i0, j0, k0, tx, ty, tz);
    if (NF == 2) fwd1[id] = gather_lerp<T>(g, src1, i0, j0, k0, tx, ty, tz);
}
template <typename T, int NF>
__global__ __launch_bounds__(kThreads) void k_sl_forward(Dims g, T s, const T *__restrict__ u,
                                                         const T *__restrict__ v,
                                                         const T *__restrict__ w,
                                                         const T *__restrict__ src0,
                                                         const T *__restrict__ src1,
                                                         T *__restrict__ fwd0,
                                                         T *__restrict__ fwd1) {
    body_sl_forward<T, NF>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, blockIdx.z, g, s, u, v, w, src0, src1, fwd0, fwd1);
}

// MacCormack correction with the 8-cell extrema clamp (src/fluid.rs:183-228):
// out = clamp(fwd + (src - bwd)/2, min8, max8) * diss away from walls,
// fwd * diss near walls, bwd = SL(fwd, -dt) evaluated at this cell only.
template <typename T>
__device__ __forceinline__ T mac_correct_one(const Dims &g, const T *__restrict__ src,
                                             const T *__restrict__ fwd, int id, bool near_wall,
                                             int i0, int j0, int k0, int fi0, int fj0, int fk0,
                                             T ftx, T fty, T ftz, T diss, bool clamp0) {
    T val;
    const T f = fwd[id];
    if (near_wall) {
        val = f;
    } else {
        const T bwd = gather_lerp<T>(g, fwd, fi0, fj0, fk0, ftx, fty, ftz);
        const T mc = f + T(0.5) * (src[id] - bwd);
        T lo, hi;
        gather_minmax<T>(g, src, i0, j0, k0, lo, hi);
        val = mc;
        if (val < lo) val = lo;
        if (val > hi) val = hi;
    }
    T o = val * diss;
    if (clamp0) o = fmax(o, T(0));
    return o;
}

template <typename T, int NF>
__device__ __forceinline__ void body_mac_correct(int i, int j, int k, 
    Dims g, T s, const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ src0, const T *__restrict__ fwd0, T *__restrict__ out0, T diss0,
    const T *__restrict__ src1, const T *__restrict__ fwd1, T *__restrict__ out1, T diss1,
    int clamp0) {
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    if (!is_interior(g, i, j, k)) {
        out0[id] = T(0);
        if (NF == 2) out1[id] = T(0);
        return;
    }
    const T su = s * u[id], sv = s * v[id], sw = s * w[id];
    int i0, j0, k0, fi0, fj0, fk0;
    T tx, ty, tz, ftx, fty, ftz;
    bool nwx, nwy, nwz, dummy;
    axis_loc(i, su, g.nx, i0, tx, nwx);
    axis_loc(j, sv, g.ny, j0, ty, nwy);
    axis_loc(k, sw, g.nz, k0, tz, nwz);
    const bool near_wall = nwx || nwy || nwz;
    // forward trace for bwd = SL(fwd, -dt): position i + s*u
    axis_loc(i, -su, g.nx, fi0, ftx, dummy);
    axis_loc(j, -sv, g.ny, fj0, fty, dummy);
    axis_loc(k, -sw, g.nz, fk0, ftz, dummy);
    out0[id] = mac_correct_one<T>(g, src0, fwd0, id, near_wall, i0, j0, k0, fi0, fj0, fk0, ftx,
                                  fty, ftz, diss0, clamp0 != 0);
    if (NF == 2)
        out1[id] = mac_correct_one<T>(g, src1, fwd1, id, near_wall, i0, j0, k0, fi0, fj0, fk0,
                                      ftx, fty, ftz, diss1, clamp0 != 0);
}
template <typename T, int NF>
__global__ __launch_bounds__(kThreads) void k_mac_correct(
    Dims g, T s, const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ src0, const T *__restrict__ fwd0, T *__restrict__ out0, T diss0,
    const T *__restrict__ src1, const T *__restrict__ fwd1, T *__restrict__ out1, T diss1,
    int clamp0) {
    body_mac_correct<T, NF>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, blockIdx.z, g, s, u, v, w, src0, fwd0, out0, diss0, src1, fwd1, out1, diss1, clamp0);
}

// MacCormack correction of velocity component `kind` (0: u, 1: v) fused with
// the forward semi-Lagrangian pass of the next component.  The reference
// advects u, v, w in place one after the other (src/fluid.rs:599-601), so the
// next component's trace runs along the velocity with this component already
// corrected; at its own cell that value is the one this thread just formed,
// and the trace reads no other cell's velocity (src/fluid.rs:160).  u, v, w
// are the velocity of this correction (for kind 1, u is already corrected).
// Same arithmetic as k_mac_correct<T, 1> followed by k_sl_forward<T, 1>.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_mac_correct_forward(
    Dims g, T s, const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ src, const T *__restrict__ fwd, T *__restrict__ out, T diss, int kind,
    const T *__restrict__ src_next, T *__restrict__ fwd_next) {
    const int i = blockIdx.x * kBX + threadIdx.x, j = blockIdx.y * kBY + threadIdx.y,
              k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    if (!is_interior(g, i, j, k)) {
        out[id] = T(0);
        fwd_next[id] = T(0);
        return;
    }
    const T su = s * u[id], sv = s * v[id], sw = s * w[id];
    int i0, j0, k0, fi0, fj0, fk0;
    T tx, ty, tz, ftx, fty, ftz;
    bool nwx, nwy, nwz, dummy;
    axis_loc(i, su, g.nx, i0, tx, nwx);
    axis_loc(j, sv, g.ny, j0, ty, nwy);
    axis_loc(k, sw, g.nz, k0, tz, nwz);
    axis_loc(i, -su, g.nx, fi0, ftx, dummy);
    axis_loc(j, -sv, g.ny, fj0, fty, dummy);
    axis_loc(k, -sw, g.nz, fk0, ftz, dummy);
    const T val = mac_correct_one<T>(g, src, fwd, id, nwx || nwy || nwz, i0, j0, k0, fi0, fj0,
                                     fk0, ftx, fty, ftz, diss, false);
    out[id] = val;
    // the next component's trace: this component replaced by its new value
    const T su2 = kind == 0 ? s * val : su, sv2 = kind == 1 ? s * val : sv;
    int n0, m0, l0;
    T nx_, ny_, nz_;
    bool nw;
    axis_loc(i, su2, g.nx, n0, nx_, nw);
    axis_loc(j, sv2, g.ny, m0, ny_, nw);
    axis_loc(k, sw, g.nz, l0, nz_, nw);
    fwd_next[id] = gather_lerp<T>(g, src_next, n0, m0, l0, nx_, ny_, nz_);
}

// ---------------------------------------------------------------------------
// Walls, vorticity confinement, buoyancy (src/fluid.rs:236-330)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void body_enforce_bc(int i, int j, int k, Dims g, T *__restrict__ u,
                                                         T *__restrict__ v, T *__restrict__ w) {
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    if (bc_u(g, i, j, k)) u[id] = T(0);
    if (bc_v(g, i, j, k)) v[id] = T(0);
    if (bc_w(g, i, j, k)) w[id] = T(0);
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_enforce_bc(Dims g, T *__restrict__ u,
                                                         T *__restrict__ v, T *__restrict__ w) {
    body_enforce_bc<T>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, blockIdx.z, g, u, v, w);
}

template <typename T>
__device__ __forceinline__ void curl_at(const Dims &g, const T *__restrict__ u,
                                        const T *__restrict__ v, const T *__restrict__ w, int id,
                                        T inv_2h, T &ox, T &oy, T &oz) {
    // src/fluid.rs:265-274
    const T dw_dy = (w[id + g.nx] - w[id - g.nx]) * inv_2h;
    const T dv_dz = (v[id + g.sxy] - v[id - g.sxy]) * inv_2h;
    const T du_dz = (u[id + g.sxy] - u[id - g.sxy]) * inv_2h;
    const T dw_dx = (w[id + 1] - w[id - 1]) * inv_2h;
    const T dv_dx = (v[id + 1] - v[id - 1]) * inv_2h;
    const T du_dy = (u[id + g.nx] - u[id - g.nx]) * inv_2h;
    ox = dw_dy - dv_dz;
    oy = du_dz - dw_dx;
    oz = dv_dx - du_dy;
}

// A sphere source applied inside another grid pass: the interior cells with
// d^2 < r^2 in cell-index coordinates get density and temperature set (as
// k_sphere, whose box covers the same interior cells).
struct SphereSrc {
    int on;
    double cx, cy, cz, r2, rho, temp;
};

template <typename T>
__device__ __forceinline__ void body_curl_mag(int i, int j, int k, Dims g, T inv_2h, const T *__restrict__ u,
                                                       const T *__restrict__ v,
                                                       const T *__restrict__ w,
                                                       T *__restrict__ cmag, const SphereSrc &sph,
                                                       T *__restrict__ rho, T *__restrict__ temp) {
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    if (!is_interior(g, i, j, k)) {
        cmag[id] = T(0);
        return;
    }
    if (sph.on) {
        const double dx = (double)i - sph.cx, dy = (double)j - sph.cy, dz = (double)k - sph.cz;
        if (dx * dx + dy * dy + dz * dz < sph.r2) {
            rho[id] = (T)sph.rho;
            temp[id] = (T)sph.temp;
        }
    }
    T ox, oy, oz;
    curl_at<T>(g, u, v, w, id, inv_2h, ox, oy, oz);
    cmag[id] = sqrt(ox * ox + oy * oy + oz * oz);
}
// |curl u| (src/fluid.rs:259-281); with sph.on, also the sphere source that
// the step applies just before (the curl reads no density or temperature).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_curl_mag(Dims g, T inv_2h, const T *__restrict__ u,
                                                       const T *__restrict__ v,
                                                       const T *__restrict__ w,
                                                       T *__restrict__ cmag, SphereSrc sph,
                                                       T *__restrict__ rho, T *__restrict__ temp) {
    body_curl_mag<T>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, blockIdx.z, g, inv_2h, u, v, w, cmag, sph, rho, temp);
}

// Confinement force eps*h*(N x w) on cells [2, n-3], then Boussinesq lift on
// interior cells, then the wall masks; writes new velocity arrays.
template <typename T>
__device__ __forceinline__ void body_confine_buoy_bc(int i, int j, int k, 
    Dims g, T dt, T inv_2h, T eps_h, T beta, T alpha, const T *__restrict__ u,
    const T *__restrict__ v, const T *__restrict__ w, const T *__restrict__ cmag,
    const T *__restrict__ rho, const T *__restrict__ temp, T *__restrict__ uo,
    T *__restrict__ vo, T *__restrict__ wo) {
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    T uu = u[id], vv = v[id], ww = w[id];
    if (i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k >= 2 && k <= g.nz - 3) {
        const T gx = (cmag[id + 1] - cmag[id - 1]) * inv_2h;
        const T gy = (cmag[id + g.nx] - cmag[id - g.nx]) * inv_2h;
        const T gz = (cmag[id + g.sxy] - cmag[id - g.sxy]) * inv_2h;
        const T glen = sqrt(gx * gx + gy * gy + gz * gz) + T(1e-10);
        const T nxv = gx / glen, nyv = gy / glen, nzv = gz / glen;
        T ox, oy, oz;
        curl_at<T>(g, u, v, w, id, inv_2h, ox, oy, oz);
        const T fx = eps_h * (nyv * oz - nzv * oy);
        const T fy = eps_h * (nzv * ox - nxv * oz);
        const T fz = eps_h * (nxv * oy - nyv * ox);
        uu += dt * fx;
        vv += dt * fy;
        ww += dt * fz;
    }
    if (is_interior(g, i, j, k)) {
        const T lift = beta * temp[id] - alpha * rho[id];
        vv += dt * lift;
    }
    if (bc_u(g, i, j, k)) uu = T(0);
    if (bc_v(g, i, j, k)) vv = T(0);
    if (bc_w(g, i, j, k)) ww = T(0);
    uo[id] = uu;
    vo[id] = vv;
    wo[id] = ww;
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_confine_buoy_bc(
    Dims g, T dt, T inv_2h, T eps_h, T beta, T alpha, const T *__restrict__ u,
    const T *__restrict__ v, const T *__restrict__ w, const T *__restrict__ cmag,
    const T *__restrict__ rho, const T *__restrict__ temp, T *__restrict__ uo,
    T *__restrict__ vo, T *__restrict__ wo) {
    body_confine_buoy_bc<T>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, blockIdx.z, g, dt, inv_2h, eps_h, beta, alpha, u, v, w, cmag, rho, temp, uo, vo, wo);
}

// ---------------------------------------------------------------------------
// Sources (src/fluid.rs:532-595 and the sphere extension).  Evaluated in
// FP64 for both storage types.
// ---------------------------------------------------------------------------
struct JetParams {
    double jet[2][6];  // x, y, z, vx, vy, vz per jet
    double pulse, r2, dt;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void k_jets(Dims g, JetParams jp, int i_lo, int j_lo,
                                                   int k_lo, int ni, int nj, int nk,
                                                   T *__restrict__ u, T *__restrict__ v,
                                                   T *__restrict__ w, T *__restrict__ rho,
                                                   T *__restrict__ temp) {
    const int a = blockIdx.x * kThreads + tid_lin();
    if (a >= ni * nj * nk) return;
    const int i = i_lo + a % ni, j = j_lo + (a / ni) % nj, k = k_lo + a / (ni * nj);
    const int id = (k * g.ny + j) * g.nx + i;
    double dd = (double)rho[id], tt = (double)temp[id], uu = (double)u[id], vv = (double)v[id],
           ww = (double)w[id];
    bool hit = false;
    for (int q = 0; q < 2; ++q) {  // jet order of src/fluid.rs:565
        const double dx = (double)i - jp.jet[q][0];
        const double dy = (double)j - jp.jet[q][1];
        const double dz = (double)k - jp.jet[q][2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < jp.r2) {
            hit = true;
            const double wf = exp(-2.2 * d2 / jp.r2);
            dd = fmin(dd + jp.dt * 9.5 * jp.pulse * wf, 3.8);
            tt = fmin(tt + jp.dt * 8.2 * jp.pulse * wf, 4.5);
            uu += jp.dt * jp.jet[q][3] * wf * 4.2;
            vv += jp.dt * jp.jet[q][4] * wf * 4.2;
            ww += jp.dt * jp.jet[q][5] * wf * 4.2;
        }
    }
    if (hit) {
        rho[id] = (T)dd;
        temp[id] = (T)tt;
        u[id] = (T)uu;
        v[id] = (T)vv;
        w[id] = (T)ww;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_sphere(Dims g, double cx, double cy, double cz,
                                                     double r2, double rho_val, double t_val,
                                                     int i_lo, int j_lo, int k_lo, int ni, int nj,
                                                     int nk, T *__restrict__ rho,
                                                     T *__restrict__ temp) {
    const int a = blockIdx.x * kThreads + tid_lin();
    if (a >= ni * nj * nk) return;
    const int i = i_lo + a % ni, j = j_lo + (a / ni) % nj, k = k_lo + a / (ni * nj);
    const double dx = (double)i - cx, dy = (double)j - cy, dz = (double)k - cz;
    if (dx * dx + dy * dy + dz * dz < r2) {
        const int id = (k * g.ny + j) * g.nx + i;
        rho[id] = (T)rho_val;
        temp[id] = (T)t_val;
    }
}

// ---------------------------------------------------------------------------
// Projection (src/fluid.rs:333-529)
// ---------------------------------------------------------------------------

// Velocity read through the wall mask: the value enforce_wall_velocity_bc
// would leave (src/fluid.rs:497 precedes the divergence).
template <typename T>
__device__ __forceinline__ void div_terms(const Dims &g, const T *__restrict__ u,
                                          const T *__restrict__ v, const T *__restrict__ w,
                                          int i, int j, int k, int id, bool masked, T &du, T &dv,
                                          T &dw) {
    T u1 = u[id + 1], u0 = u[id], v1 = v[id + g.nx], v0 = v[id], w1 = w[id + g.sxy], w0 = w[id];
    if (masked) {
        if (bc_u(g, i + 1, j, k)) u1 = T(0);
        if (bc_u(g, i, j, k)) u0 = T(0);
        if (bc_v(g, i, j + 1, k)) v1 = T(0);
        if (bc_v(g, i, j, k)) v0 = T(0);
        if (bc_w(g, i, j, k + 1)) w1 = T(0);
        if (bc_w(g, i, j, k)) w0 = T(0);
    }
    du = u1 - u0;
    dv = v1 - v0;
    dw = w1 - w0;
}

// A x at an interior cell (src/fluid.rs:356-394): pure-Neumann 7-point
// operator, a neighbour counts only if it is interior.  X(idx) supplies x.
template <typename T, typename X>
__device__ __forceinline__ T lap_at(const Dims &g, int i, int j, int k, int id, X x) {
    if constexpr (sizeof(T) == 8) {
        T diag = T(0), nb = T(0);
        if (i > 1) { diag += T(1); nb += x(id - 1); }
        if (i + 2 < g.nx) { diag += T(1); nb += x(id + 1); }
        if (j > 1) { diag += T(1); nb += x(id - g.nx); }
        if (j + 2 < g.ny) { diag += T(1); nb += x(id + g.nx); }
        if (k > 1) { diag += T(1); nb += x(id - g.sxy); }
        if (k + 2 < g.nz) { diag += T(1); nb += x(id + g.sxy); }
        return diag * x(id) - nb;
    } else {
        const T xc = x(id);
        T acc = T(0);
        if (i > 1) acc += xc - x(id - 1);
        if (i + 2 < g.nx) acc += xc - x(id + 1);
        if (j > 1) acc += xc - x(id - g.nx);
        if (j + 2 < g.ny) acc += xc - x(id + g.nx);
        if (k > 1) acc += xc - x(id - g.sxy);
        if (k + 2 < g.nz) acc += xc - x(id + g.sxy);
        return acc;
    }
}

// b = -h*div(u) (src/fluid.rs:346) and, for the norm, sum of div^2
// (src/fluid.rs:344-347); one partial per workgroup.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_divergence(Dims g, int zchunk, T h, T inv_h,
                                                         int masked, const T *__restrict__ u,
                                                         const T *__restrict__ v,
                                                         const T *__restrict__ w,
                                                         T *__restrict__ b,
                                                         double *__restrict__ part) {
    __shared__ double lds[8];
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int kb = blockIdx.z * zchunk;
    const int ke = min(kb + zchunk, g.nz);
    double acc[1] = {0.0};
    if (i < g.nx && j < g.ny) {
        for (int k = kb; k < ke; ++k) {
            const int id = (k * g.ny + j) * g.nx + i;
            T bv = T(0);
            if (is_interior(g, i, j, k)) {
                T du, dv, dw;
                div_terms<T>(g, u, v, w, i, j, k, id, masked != 0, du, dv, dw);
                const T dsum = du + dv + dw;
                const double dval = (double)(dsum * inv_h);
                bv = -h * dsum;
                acc[0] += dval * dval;
            }
            b[id] = bv;
        }
    }
    block_sum<1>(acc, lds);
    if (tid_lin() == 0)
        part[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = acc[0];
}

// max |div u| over interior cells, one partial per workgroup.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_div_max(Dims g, int zchunk, T inv_h,
                                                      const T *__restrict__ u,
                                                      const T *__restrict__ v,
                                                      const T *__restrict__ w,
                                                      double *__restrict__ part) {
    __shared__ double lds[8];
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int kb = blockIdx.z * zchunk;
    const int ke = min(kb + zchunk, g.nz);
    double m = 0.0;
    if (i < g.nx && j < g.ny) {
        for (int k = kb; k < ke; ++k) {
            if (!is_interior(g, i, j, k)) continue;
            const int id = (k * g.ny + j) * g.nx + i;
            T du, dv, dw;
            div_terms<T>(g, u, v, w, i, j, k, id, false, du, dv, dw);
            m = fmax(m, fabs((double)((du + dv + dw) * inv_h)));
        }
    }
    m = block_max(m, lds);
    if (tid_lin() == 0)
        part[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = m;
}

// ---------------------------------------------------------------------------
// CG kernels.  Workgroup = 64 x 4 columns of cells; every thread owns ZU
// consecutive z cells of its column.  All loads of a thread are issued before
// the workgroup combines the previous kernel's partial sums (a barrier on
// gfx950 does not wait for outstanding global loads), so the fixed-order
// reduction overlaps the memory traffic.  Workgroups are numbered so that
// each of the 8 XCDs receives one contiguous slab of tiles (block b runs on
// XCD b % 8), which keeps y/z halo reads inside one L2.
// ---------------------------------------------------------------------------
struct Tiling {
    int gx, gy, gz;  // tiles of 64 x, 4 y and ZU z cells
    int nb;          // gx * gy * gz workgroups (= partial sums)
    int remap;       // 1 when nb % 8 == 0: XCD-contiguous tile order
};

__device__ __forceinline__ void tile_coords(const Tiling &t, int &bx, int &by, int &bz) {
    int b = blockIdx.x;
    if (t.remap) b = (b & 7) * (t.nb >> 3) + (b >> 3);
    bx = b % t.gx;
    const int rest = b / t.gx;
    by = rest % t.gy;
    bz = rest / t.gy;
}

// A x at an interior cell from its 7 stencil values (src/fluid.rs:356-394);
// v* say whether each neighbour is an interior cell.
template <typename T>
__device__ __forceinline__ T lap7(T xc, T xm, T xp, T ym, T yp, T zm, T zp, bool vxm, bool vxp,
                                  bool vym, bool vyp, bool vzm, bool vzp) {
    if constexpr (sizeof(T) == 8) {
        T diag = T(0), nb = T(0);
        if (vxm) { diag += T(1); nb += xm; }
        if (vxp) { diag += T(1); nb += xp; }
        if (vym) { diag += T(1); nb += ym; }
        if (vyp) { diag += T(1); nb += yp; }
        if (vzm) { diag += T(1); nb += zm; }
        if (vzp) { diag += T(1); nb += zp; }
        return diag * xc - nb;
    } else {
        T acc = T(0);
        if (vxm) acc += xc - xm;
        if (vxp) acc += xc - xp;
        if (vym) acc += xc - ym;
        if (vyp) acc += xc - yp;
        if (vzm) acc += xc - zm;
        if (vzp) acc += xc - zp;
        return acc;
    }
}

// Store for CG vectors: a plain store (non-temporal stores measured ~2%
// slower at 128^3, as the next launch re-reads the vector from L2/MALL).
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) { *p = v; }

// Column geometry shared by the CG kernels.
struct Col {
    int i, j, k0;
    int ic, jc;          // clamped column (always a valid address)
    int xm, xp, ym, yp;  // clamped in-plane neighbour offsets from (ic, jc)
    bool valid;
};

template <int ZU>
__device__ __forceinline__ Col col_at(const Dims &g, int bx, int by, int bz) {
    Col c;
    c.i = bx * kBX + threadIdx.x;
    c.j = by * kBY + threadIdx.y;
    c.k0 = bz * ZU;
    c.valid = c.i < g.nx && c.j < g.ny;
    c.ic = min(c.i, g.nx - 1);
    c.jc = min(c.j, g.ny - 1);
    c.xm = (c.ic > 0 ? -1 : 0);
    c.xp = (c.ic < g.nx - 1 ? 1 : 0);
    c.ym = (c.jc > 0 ? -g.nx : 0);
    c.yp = (c.jc < g.ny - 1 ? g.nx : 0);
    return c;
}

template <int ZU>
__device__ __forceinline__ Col col_of(const Dims &g, const Tiling &tl) {
    int bx, by, bz;
    tile_coords(tl, bx, by, bz);
    return col_at<ZU>(g, bx, by, bz);
}

__device__ __forceinline__ int col_id(const Dims &g, const Col &c, int k) {
    k = min(max(k, 0), g.nz - 1);
    return (k * g.ny + c.jc) * g.nx + c.ic;
}

// Divergence + CG initialisation (src/fluid.rs:497-498, 450-460):
// b = -h div(masked u) (or the stored b), r = b - A p; partials of r.r, b.b.
template <typename T, int ZU>
__device__ __forceinline__ void body_div_cginit(const Col &c, int slot, Dims g, T h, int with_div,
                                                const T *__restrict__ u, const T *__restrict__ v,
                                                const T *__restrict__ w, const T *__restrict__ p,
                                                T *__restrict__ b, T *__restrict__ r,
                                                double *__restrict__ part_rr,
                                                double *__restrict__ part_bb, CgState *st) {
    __shared__ double lds[16];
    double acc[2] = {0.0, 0.0};
    if (c.valid) {
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            const int k = c.k0 + a;
            if (k >= g.nz) break;
            const int id = (k * g.ny + c.j) * g.nx + c.i;
            T bv = T(0), rv = T(0);
            if (is_interior(g, c.i, c.j, k)) {
                if (with_div) {
                    T du, dv, dw;
                    div_terms<T>(g, u, v, w, c.i, c.j, k, id, true, du, dv, dw);
                    bv = -h * (du + dv + dw);
                } else {
                    bv = b[id];
                }
                const T q = lap7<T>(p[id], p[id - 1], p[id + 1], p[id - g.nx], p[id + g.nx],
                                    p[id - g.sxy], p[id + g.sxy], c.i > 1, c.i + 2 < g.nx,
                                    c.j > 1, c.j + 2 < g.ny, k > 1, k + 2 < g.nz);
                rv = bv - q;
            }
            if (with_div) b[id] = bv;
            r[id] = rv;
            acc[0] += (double)rv * (double)rv;
            acc[1] += (double)bv * (double)bv;
        }
    }
    block_sum<2>(acc, lds);
    if (tid_lin() == 0) {
        part_rr[slot] = acc[0];
        part_bb[slot] = acc[1];
        if (slot == 0) {
            st->done = 0;
            st->iters = 0;
            st->converged = 0;
        }
    }
}

template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void k_div_cginit(Dims g, Tiling tl, T h, int with_div,
                                                         const T *__restrict__ u,
                                                         const T *__restrict__ v,
                                                         const T *__restrict__ w,
                                                         const T *__restrict__ p,
                                                         T *__restrict__ b, T *__restrict__ r,
                                                         double *__restrict__ part_rr,
                                                         double *__restrict__ part_bb,
                                                         CgState *st) {
    body_div_cginit<T, ZU>(col_of<ZU>(g, tl), blockIdx.x, g, h, with_div, u, v, w, p, b, r,
                           part_rr, part_bb, st);
}

// CG kernel 1 of iteration `it` (src/fluid.rs:463-469, 482-486): finish
// r.r of the previous iteration, apply the stopping rules, form
// d = r + beta d_old (d = r at it = 0) on the fly, q = A d, partial d.q.
// d is double-buffered so that neighbouring columns recompute it.
template <typename T, int ZU, bool FIRST>
__device__ __forceinline__ void body_cg_dir(const Col &c, int slot, int nb, Dims g, int it,
                                            double rel_tol2, const double *__restrict__ part_rr,
                                            const double *__restrict__ part_bb, CgState *st,
                                            const T *__restrict__ r, const T *__restrict__ dold,
                                            T *__restrict__ dnew, double *__restrict__ part_dq,
                                            int *done_mirror, int solve_id) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    constexpr bool first = FIRST;
    // 1) scalars and partial sums of the previous kernel, then every data
    //    load of this thread: one memory round trip for all of them.  The
    //    stop flag was written by an earlier kernel, so a plain load sees it.
    const int flag_reg = st->done;
    EarlyPartials prr, pbb;
    prr.load(part_rr, nb);
    if (first) pbb.load(part_bb, nb);
    const double delta_prev = first ? 0.0 : st->delta[(it - 1) & 1];
    const double b2_prev = first ? 0.0 : st->b2;
    T rz[ZU + 2], dz[ZU + 2], rn[ZU][4], dn[ZU][4];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int id = col_id(g, c, c.k0 - 1 + a);
        rz[a] = r[id];
        dz[a] = first ? T(0) : dold[id];
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int id = col_id(g, c, c.k0 + a);
        const int o[4] = {c.xm, c.xp, c.ym, c.yp};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rn[a][q] = r[id + o[q]];
            dn[a][q] = first ? T(0) : dold[id + o[q]];
        }
    }
    // 2) stopping rules (every workgroup reaches the same decision); thread
    //    0's flag is broadcast so that all waves take the same branch
    const bool lead = slot == 0 && tid_lin() == 0;
    double beta = 0.0;
    if (tid_lin() == 0) flag_lds[0] = flag_reg;
    const double rr = prr.total(part_rr, nb, lds);  // its barrier publishes flag_lds
    if (flag_lds[0]) return;
    if (first) {
        const double bb = pbb.total(part_bb, nb, lds);
        const bool stop = !(rr > 1e-24) || (rel_tol2 > 0.0 && rr <= rel_tol2 * bb);
        if (lead) {
            st->b2 = bb;
            st->r2_init = rr;
            st->delta[0] = rr;
            if (stop) {
                st->done = 1;
                st->iters = 0;
                st->r2_final = rr;
                st->converged = 1;
                *done_mirror = solve_id;
            }
        }
        if (stop) return;
    } else {
        const bool stop = rr < 1e-20 || (rel_tol2 > 0.0 && rr <= rel_tol2 * b2_prev);
        if (lead) {
            if (stop) {
                st->done = 1;
                st->iters = it;
                st->r2_final = rr;
                st->converged = 1;
                *done_mirror = solve_id;
            } else {
                st->delta[it & 1] = rr;
            }
        }
        if (stop) return;
        beta = rr / delta_prev;
    }
    // 3) d, A d, d.(A d)
    const T tb = (T)beta;
    double acc[1] = {0.0};
    T dcol[ZU + 2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) dcol[a] = first ? rz[a] : rz[a] + tb * dz[a];
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        if (c.valid && k < g.nz) {
            const int id = (k * g.ny + c.j) * g.nx + c.i;
            T x4[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) x4[q] = first ? rn[a][q] : rn[a][q] + tb * dn[a][q];
            const T xc = dcol[a + 1];
            if (is_interior(g, c.i, c.j, k)) {
                const T q = lap7<T>(xc, x4[0], x4[1], x4[2], x4[3], dcol[a], dcol[a + 2],
                                    c.i > 1, c.i + 2 < g.nx, c.j > 1, c.j + 2 < g.ny, k > 1,
                                    k + 2 < g.nz);
                acc[0] += (double)xc * (double)q;
            }
            st_stream(dnew + id, xc);
        }
    }
    block_sum<1>(acc, lds);
    if (tid_lin() == 0) part_dq[slot] = acc[0];
}

template <typename T, int ZU, bool FIRST>
__global__ __launch_bounds__(kThreads) void k_cg_dir(Dims g, Tiling tl, int it, double rel_tol2,
                                                     const double *__restrict__ part_rr,
                                                     const double *__restrict__ part_bb,
                                                     CgState *st, const T *__restrict__ r,
                                                     const T *__restrict__ dold,
                                                     T *__restrict__ dnew,
                                                     double *__restrict__ part_dq,
                                                     int *done_mirror, int solve_id) {
    body_cg_dir<T, ZU, FIRST>(col_of<ZU>(g, tl), blockIdx.x, tl.nb, g, it, rel_tol2, part_rr,
                              part_bb, st, r, dold, dnew, part_dq, done_mirror, solve_id);
}

// CG kernel 2 of iteration `it` (src/fluid.rs:466-478): finish d.q,
// alpha, p += alpha d, r -= alpha A d (A d recomputed), partial r.r.
template <typename T, int ZU>
__device__ __forceinline__ void body_cg_update(const Col &c, int slot, int nb, Dims g, int it,
                                               const double *__restrict__ part_dq, CgState *st,
                                               const T *__restrict__ d, T *__restrict__ p,
                                               T *__restrict__ r, double *__restrict__ part_rr,
                                               int *done_mirror, int solve_id) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    const int flag_reg = st->done;
    EarlyPartials pdq;
    pdq.load(part_dq, nb);
    const double delta = st->delta[it & 1];
    T dz[ZU + 2], dn[ZU][4], pc[ZU], rc[ZU];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) dz[a] = d[col_id(g, c, c.k0 - 1 + a)];
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int id = col_id(g, c, c.k0 + a);
        dn[a][0] = d[id + c.xm];
        dn[a][1] = d[id + c.xp];
        dn[a][2] = d[id + c.ym];
        dn[a][3] = d[id + c.yp];
        pc[a] = p[id];
        rc[a] = r[id];
    }
    const bool lead = slot == 0 && tid_lin() == 0;
    if (tid_lin() == 0) flag_lds[0] = flag_reg;
    const double dq = pdq.total(part_dq, nb, lds);
    if (flag_lds[0]) return;
    if (fabs(dq) < 1e-28) {
        if (lead) {
            st->done = 1;
            st->iters = it;
            st->converged = 1;
            st->r2_final = delta;
            *done_mirror = solve_id;
        }
        return;
    }
    const T ta = (T)(delta / dq);
    double acc[1] = {0.0};
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        if (c.valid && k < g.nz) {
            const int id = (k * g.ny + c.j) * g.nx + c.i;
            const T dc = dz[a + 1];
            T q = T(0);
            if (is_interior(g, c.i, c.j, k))
                q = lap7<T>(dc, dn[a][0], dn[a][1], dn[a][2], dn[a][3], dz[a], dz[a + 2],
                            c.i > 1, c.i + 2 < g.nx, c.j > 1, c.j + 2 < g.ny, k > 1,
                            k + 2 < g.nz);
            st_stream(p + id, pc[a] + ta * dc);
            const T rn = rc[a] - ta * q;
            st_stream(r + id, rn);
            acc[0] += (double)rn * (double)rn;
        }
    }
    block_sum<1>(acc, lds);
    if (tid_lin() == 0) part_rr[slot] = acc[0];
}

template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void k_cg_update(Dims g, Tiling tl, int it,
                                                        const double *__restrict__ part_dq,
                                                        CgState *st, const T *__restrict__ d,
                                                        T *__restrict__ p, T *__restrict__ r,
                                                        double *__restrict__ part_rr,
                                                        int *done_mirror, int solve_id) {
    body_cg_update<T, ZU>(col_of<ZU>(g, tl), blockIdx.x, tl.nb, g, it, part_dq, st, d, p, r,
                          part_rr, done_mirror, solve_id);
}

// Closes a solve that ran out of iterations (or had max_iters == 0).
__global__ __launch_bounds__(kThreads) void k_cg_finalize(int max_iters, double rel_tol2,
                                                          const double *__restrict__ part_rr,
                                                          const double *__restrict__ part_bb,
                                                          int nb, CgState *st,
                                                          int *__restrict__ iters_log,
                                                          unsigned long long *iters_mirror,
                                                          int solve_id) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    if (!block_flag(&st->done, flag_lds)) {
        const double rr = sum_partials(part_rr, nb, lds);
        double bb = 0.0;
        if (max_iters == 0) bb = sum_partials(part_bb, nb, lds);
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
    if (iters_log && tid_lin() == 0) *iters_log = st->iters;
    // the host queues the next solve at once when a recent one ran to the
    // cap; id and count go out in one 64-bit store, so they never tear
    if (iters_mirror && tid_lin() == 0)
        *iters_mirror = ((unsigned long long)(unsigned)solve_id << 32) | (unsigned)st->iters;
}

// Red-black SOR sweep of one colour (src/fluid.rs:400-437).  Same-colour
// cells are independent, so this equals the reference's sequential loop.
// (b + sum of neighbours) / diag of the Gauss-Seidel update
// (src/fluid.rs:434).  FP64 divides, as the reference does.  FP32
// multiplies by the hardware reciprocal of diag (an integer 1..6): within two
// ulps of the quotient, for ~2 instructions instead of an IEEE division's ~10,
// which made up a quarter of the FP32 sweep's vector instructions.
__device__ __forceinline__ double sor_quot(double x, double diag) { return x / diag; }
__device__ __forceinline__ float sor_quot(float x, float diag) {
    return x * __builtin_amdgcn_rcpf(diag);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_rbgs(Dims g, int color, T omega,
                                                   const T *__restrict__ b, T *__restrict__ p) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    if (!is_interior(g, i, j, k) || ((i + j + k) & 1) != color) return;
    const int id = (k * g.ny + j) * g.nx + i;
    T diag = T(0), nb = T(0);
    if (i > 1) { diag += T(1); nb += p[id - 1]; }
    if (i + 2 < g.nx) { diag += T(1); nb += p[id + 1]; }
    if (j > 1) { diag += T(1); nb += p[id - g.nx]; }
    if (j + 2 < g.ny) { diag += T(1); nb += p[id + g.nx]; }
    if (k > 1) { diag += T(1); nb += p[id - g.sxy]; }
    if (k + 2 < g.nz) { diag += T(1); nb += p[id + g.sxy]; }
    const T gs = sor_quot(b[id] + nb, diag);
    p[id] = (T(1) - omega) * p[id] + omega * gs;
}

// Gradient subtraction and the closing wall masks (src/fluid.rs:509-528):
// the projected velocity of cell id.
template <typename T>
__device__ __forceinline__ void gradient_bc_at(const Dims &g, int i, int j, int k, int id,
                                               T inv_h, const T *__restrict__ p,
                                               const T *__restrict__ u, const T *__restrict__ v,
                                               const T *__restrict__ w, T &uu, T &vv, T &ww) {
    uu = u[id];
    vv = v[id];
    ww = w[id];
    if (is_interior(g, i, j, k)) {
        const T pc = p[id];
        if (i > 1) uu -= (pc - p[id - 1]) * inv_h;
        if (j > 1) vv -= (pc - p[id - g.nx]) * inv_h;
        if (k > 1) ww -= (pc - p[id - g.sxy]) * inv_h;
    }
    if (bc_u(g, i, j, k)) uu = T(0);
    if (bc_v(g, i, j, k)) vv = T(0);
    if (bc_w(g, i, j, k)) ww = T(0);
}
template <typename T>
__device__ __forceinline__ void body_gradient_bc(int i, int j, int k, Dims g, T inv_h,
                                                          const T *__restrict__ p,
                                                          T *__restrict__ u, T *__restrict__ v,
                                                          T *__restrict__ w) {
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    T uu, vv, ww;
    gradient_bc_at<T>(g, i, j, k, id, inv_h, p, u, v, w, uu, vv, ww);
    u[id] = uu;
    v[id] = vv;
    w[id] = ww;
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gradient_bc(Dims g, T inv_h,
                                                          const T *__restrict__ p,
                                                          T *__restrict__ u, T *__restrict__ v,
                                                          T *__restrict__ w) {
    body_gradient_bc<T>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, blockIdx.z, g, inv_h, p, u, v, w);
}

// The end of a step's projection fused with the start of its scalar
// advection: the projected velocity of the cell (as k_gradient_bc), then the
// forward semi-Lagrangian pass of density and temperature along it (as
// k_sl_forward<T, 2>).  The back-trace reads only the cell's own velocity
// (src/fluid.rs:160), so one pass serves both, with the same arithmetic.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gradient_sl_forward(
    Dims g, T inv_h, const T *__restrict__ p, T *__restrict__ u, T *__restrict__ v,
    T *__restrict__ w, T s, const T *__restrict__ src0, const T *__restrict__ src1,
    T *__restrict__ fwd0, T *__restrict__ fwd1) {
    const int i = blockIdx.x * kBX + threadIdx.x, j = blockIdx.y * kBY + threadIdx.y,
              k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    T uu, vv, ww;
    gradient_bc_at<T>(g, i, j, k, id, inv_h, p, u, v, w, uu, vv, ww);
    u[id] = uu;
    v[id] = vv;
    w[id] = ww;
    if (!is_interior(g, i, j, k)) {
        fwd0[id] = T(0);
        fwd1[id] = T(0);
        return;
    }
    int i0, j0, k0;
    T tx, ty, tz;
    bool nw;
    axis_loc(i, s * uu, g.nx, i0, tx, nw);
    axis_loc(j, s * vv, g.ny, j0, ty, nw);
    axis_loc(k, s * ww, g.nz, k0, tz, nw);
    fwd0[id] = gather_lerp<T>(g, src0, i0, j0, k0, tx, ty, tz);
    fwd1[id] = gather_lerp<T>(g, src1, i0, j0, k0, tx, ty, tz);
}

// ---------------------------------------------------------------------------
// Mixed-precision projection: FP64 pressure and right-hand side around an
// FP32 inner CG (iterative refinement).  Velocities stay FP32.
// ---------------------------------------------------------------------------

// b64 = -h div(masked u) in FP64 (src/fluid.rs:346 with the BC of :497).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_div64(Dims g, double h, const T *__restrict__ u,
                                                    const T *__restrict__ v,
                                                    const T *__restrict__ w,
                                                    double *__restrict__ b64) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    double b = 0.0;
    if (is_interior(g, i, j, k)) {
        // differences of the stored values are formed in FP64
        const double u1 = bc_u(g, i + 1, j, k) ? 0.0 : (double)u[id + 1];
        const double u0 = bc_u(g, i, j, k) ? 0.0 : (double)u[id];
        const double v1 = bc_v(g, i, j + 1, k) ? 0.0 : (double)v[id + g.nx];
        const double v0 = bc_v(g, i, j, k) ? 0.0 : (double)v[id];
        const double w1 = bc_w(g, i, j, k + 1) ? 0.0 : (double)w[id + g.sxy];
        const double w0 = bc_w(g, i, j, k) ? 0.0 : (double)w[id];
        b = -h * ((u1 - u0) + (v1 - v0) + (w1 - w0));
    }
    b64[id] = b;
}

// r = b64 - A p64 in FP64; r32 = (float) r feeds the inner FP32 CG; one
// partial max|r| and one partial r.r per workgroup.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_residual64(Dims g, int zchunk,
                                                         const double *__restrict__ b64,
                                                         const double *__restrict__ p64,
                                                         T *__restrict__ r32,
                                                         double *__restrict__ part_max,
                                                         double *__restrict__ part_sum) {
    __shared__ double lds[8];
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int kb = blockIdx.z * zchunk;
    const int ke = min(kb + zchunk, g.nz);
    double m = 0.0, acc[1] = {0.0};
    if (i < g.nx && j < g.ny) {
        for (int k = kb; k < ke; ++k) {
            const int id = (k * g.ny + j) * g.nx + i;
            double r = 0.0;
            if (is_interior(g, i, j, k))
                r = b64[id] - lap_at<double>(g, i, j, k, id, [&](int a) { return p64[a]; });
            r32[id] = (T)r;
            m = fmax(m, fabs(r));
            acc[0] += r * r;
        }
    }
    m = block_max(m, lds);
    block_sum<1>(acc, lds);
    if (tid_lin() == 0) {
        const int bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        part_max[bid] = m;
        part_sum[bid] = acc[0];
    }
}

// p64 += e (the inner solve's correction)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_add64(const T *__restrict__ e, long n,
                                                    double *__restrict__ p64) {
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads)
        p64[a] += (double)e[a];
}

// Gradient subtraction with the FP64 pressure, rounded once to the storage
// type, then the closing wall masks (src/fluid.rs:509-528).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gradient64_bc(Dims g, double inv_h,
                                                            const double *__restrict__ p,
                                                            T *__restrict__ u,
                                                            T *__restrict__ v,
                                                            T *__restrict__ w) {
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int id = (k * g.ny + j) * g.nx + i;
    T uu = u[id], vv = v[id], ww = w[id];
    if (is_interior(g, i, j, k)) {
        const double pc = p[id];
        if (i > 1) uu = (T)((double)uu - (pc - p[id - 1]) * inv_h);
        if (j > 1) vv = (T)((double)vv - (pc - p[id - g.nx]) * inv_h);
        if (k > 1) ww = (T)((double)ww - (pc - p[id - g.sxy]) * inv_h);
    }
    if (bc_u(g, i, j, k)) uu = T(0);
    if (bc_v(g, i, j, k)) vv = T(0);
    if (bc_w(g, i, j, k)) ww = T(0);
    u[id] = uu;
    v[id] = vv;
    w[id] = ww;
}

// ---------------------------------------------------------------------------
// Telemetry (src/fluid.rs:613-672), FP64 accumulation.
// Partial layout per workgroup: ke, enstrophy, mass, cx, cy, cz, peak.
// ---------------------------------------------------------------------------
constexpr int kTelParts = 7;

template <typename T>
__device__ __forceinline__ void body_telemetry_part(int i, int j, int kb, int ke, int slot, Dims g,
                                                    double h, const T *__restrict__ u,
                                                    const T *__restrict__ v,
                                                    const T *__restrict__ w,
                                                    const T *__restrict__ rho,
                                                    double *__restrict__ part,
                                                    int *nonfinite = nullptr) {
    __shared__ double lds[32];
    const double dv = h * h * h;
    bool bad = false;
    const double inv_2h = 0.5 / h;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double peak = 0.0;
    // interior cells only (src/fluid.rs:622-624): whole columns of them, and
    // planes 1 .. nz-2 of the chunk, so every neighbour index is in the grid.
    // Marching in z, the z-neighbours of u and v rotate through registers.
    if (i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2) {
        const int k0 = max(kb, 1), k1 = min(ke, g.nz - 1);
        const double fx = 2.0 * ((double)i / (double)(g.nx - 1)) - 1.0;
        const double fy = 2.0 * ((double)j / (double)(g.ny - 1)) - 1.0;
        int id = (k0 * g.ny + j) * g.nx + i;
        double um = 0.0, uc = 0.0, vm = 0.0, vc = 0.0;
        if (k0 < k1) {
            um = u[id - g.sxy];
            uc = u[id];
            vm = v[id - g.sxy];
            vc = v[id];
        }
#pragma unroll 2
        for (int k = k0; k < k1; ++k, id += g.sxy) {
            const double up = u[id + g.sxy], vp = v[id + g.sxy], wc = w[id];
            const double u2 = uc * uc + vc * vc + wc * wc;
            acc[0] += 0.5 * u2 * dv;
            const double dw_dy = ((double)w[id + g.nx] - (double)w[id - g.nx]) * inv_2h;
            const double dv_dz = (vp - vm) * inv_2h;
            const double du_dz = (up - um) * inv_2h;
            const double dw_dx = ((double)w[id + 1] - (double)w[id - 1]) * inv_2h;
            const double dv_dx = ((double)v[id + 1] - (double)v[id - 1]) * inv_2h;
            const double du_dy = ((double)u[id + g.nx] - (double)u[id - g.nx]) * inv_2h;
            const double ox = dw_dy - dv_dz, oy = du_dz - dw_dx, oz = dv_dx - du_dy;
            const double w2 = ox * ox + oy * oy + oz * oz;
            acc[1] += 0.5 * w2 * dv;
            peak = fmax(peak, sqrt(w2));
            const double r = rho[id];
            // a non-finite velocity or density reads as a non-finite |u|^2,
            // |w|^2 or rho (inf - inf and NaN alike)
            bad |= !isfinite(u2) || !isfinite(w2) || !isfinite(r);
            // cells at or below the threshold add +0.0, which leaves the
            // sums' bits as skipping them would
            const bool mass = r > 1e-8;
            const double m = mass ? r * dv : 0.0;
            acc[2] += m;
            acc[3] += mass ? fx * m : 0.0;
            acc[4] += mass ? fy * m : 0.0;
            acc[5] += mass ? (2.0 * ((double)k / (double)(g.nz - 1)) - 1.0) * m : 0.0;
            um = uc;
            uc = up;
            vm = vc;
            vc = vp;
        }
    }
    // the device flag of src/fluid.rs's silent NaN propagation: an integer
    // OR, the same whatever the order
    if (bad && nonfinite) atomicOr(nonfinite, 1);
    block_sum<6>(acc, lds);
    peak = block_max(peak, lds);
    if (tid_lin() == 0) {
        double *o = part + (size_t)slot * kTelParts;
        for (int a = 0; a < 6; ++a) o[a] = acc[a];
        o[6] = peak;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_telemetry_part(Dims g, int zchunk, double h,
                                                             const T *__restrict__ u,
                                                             const T *__restrict__ v,
                                                             const T *__restrict__ w,
                                                             const T *__restrict__ rho,
                                                             double *__restrict__ part,
                                                             int *nonfinite) {
    const int kb = blockIdx.z * zchunk;
    body_telemetry_part<T>(blockIdx.x * kBX + threadIdx.x, blockIdx.y * kBY + threadIdx.y, kb,
                           min(kb + zchunk, g.nz),
                           (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x, g, h,
                           u, v, w, rho, part, nonfinite);
}

// One workgroup: fixed-order combination of the partials into a
// FluidTelemetryFrame {time, ke, enstrophy, peak, mass, com xyz}.
__device__ __forceinline__ void body_telemetry_final(const double *__restrict__ part,
                                                              int nb, double time_sec,
                                                              double *__restrict__ out) {
    __shared__ double lds[32];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double peak = 0.0;
    for (int a = tid_lin(); a < nb; a += kThreads) {
        const double *o = part + (size_t)a * kTelParts;
        for (int c = 0; c < 6; ++c) acc[c] += o[c];
        peak = fmax(peak, o[6]);
    }
    block_sum<6>(acc, lds);
    peak = block_max(peak, lds);
    if (tid_lin() == 0) {
        double mass = acc[2], cx = acc[3], cy = acc[4], cz = acc[5];
        if (mass > 1e-12) {
            const double inv = 1.0 / mass;  // Vec3 /= f64 (math.rs:226-239)
            cx = cx * inv;
            cy = cy * inv;
            cz = cz * inv;
        }
        out[0] = time_sec;
        out[1] = acc[0];
        out[2] = acc[1];
        out[3] = peak;
        out[4] = mass;
        out[5] = fmin(fmax(cx, -1.0), 1.0);
        out[6] = fmin(fmax(cy, -1.0), 1.0);
        out[7] = fmin(fmax(cz, -1.0), 1.0);
    }
}

__global__ __launch_bounds__(kThreads) void k_telemetry_final(const double *__restrict__ part,
                                                              int nb, double time_sec,
                                                              double *__restrict__ out) {
    body_telemetry_final(part, nb, time_sec, out);
}

// ---------------------------------------------------------------------------
// Small utilities
// ---------------------------------------------------------------------------
// One workgroup: the total of nb partials in the order of
// EarlyPartialsT<16>::total (so a kernel that reads this total gets the
// value it would have summed itself).  32 registers per thread keep every
// load of up to 8192 partials in flight at once; EarlyPartialsT<32> adds
// the same terms in the same order as <16> (its registers in index order,
// then the rest in index order).
__global__ __launch_bounds__(kThreads) void k_partials_total(const double *__restrict__ part,
                                                             int nb, double *__restrict__ out) {
    __shared__ double lds[8];
    EarlyPartialsT<32> v;
    v.load(part, nb);
    const double s = v.total(part, nb, lds);
    if (tid_lin() == 0) *out = s;
}

__global__ __launch_bounds__(kThreads) void k_partials_final(const double *__restrict__ part,
                                                             int nb, int op,
                                                             double *__restrict__ out) {
    __shared__ double lds[8];
    if (op == 0) {
        const double s = sum_partials(part, nb, lds);
        if (tid_lin() == 0) *out = s;
    } else {
        double m = 0.0;
        for (int a = tid_lin(); a < nb; a += kThreads) m = fmax(m, part[a]);
        m = block_max(m, lds);
        if (tid_lin() == 0) *out = m;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_max_part(const T *__restrict__ x, long n,
                                                       double *__restrict__ part) {
    __shared__ double lds[8];
    double m = 0.0;
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads)
        m = fmax(m, (double)x[a]);
    m = block_max(m, lds);
    if (tid_lin() == 0) part[blockIdx.x] = m;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kThreads) void k_convert(const Tin *__restrict__ in, long n,
                                                      Tout *__restrict__ out) {
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads)
        out[a] = (Tout)in[a];
}

// Plain streaming copy (16-byte vectors), the reference point for measured
// bandwidth and for scaling the memory-traffic counters.
__global__ __launch_bounds__(kThreads) void k_stream_copy(const float4 *__restrict__ in, long n4,
                                                          float4 *__restrict__ out) {
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n4; a += (long)gridDim.x * kThreads)
        out[a] = in[a];
}

// Fills x with full-entropy values in [-1, 1) (a SplitMix64 hash of the
// index): bandwidth measured on zero-filled buffers is optimistic on MI355X,
// so the calibration copy streams data of the same character as the fields.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill_hash(T *__restrict__ x, long n) {
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads) {
        unsigned long long z = (unsigned long long)a + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x[a] = (T)((double)(z >> 11) * (1.0 / 4503599627370496.0) - 1.0);
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill(T *__restrict__ x, long n, T val) {
    for (long a = (long)blockIdx.x * kThreads + tid_lin(); a < n; a += (long)gridDim.x * kThreads)
        x[a] = val;
}

// Point sampling of src/fluid.rs:94-128, evaluated in FP64.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sample_points(Dims g, const T *__restrict__ f,
                                                            const double *__restrict__ xyz,
                                                            long npts, double *__restrict__ out) {
    const long a = (long)blockIdx.x * kThreads + tid_lin();
    if (a >= npts) return;
    double c[3];
    int i0[3];
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; ++d) {
        double x = xyz[3 * a + d];
        if (x < 0.5) x = 0.5;
        if (x > (double)n[d] - 1.500001) x = (double)n[d] - 1.500001;
        const double fl = floor(x);
        i0[d] = min(max((int)fl, 0), n[d] - 2);
        c[d] = x - fl;
    }
    const int b = (i0[2] * g.ny + i0[1]) * g.nx + i0[0];
    out[a] = lerp3<double>(f[b], f[b + 1], f[b + g.nx], f[b + g.nx + 1], f[b + g.sxy],
                           f[b + g.sxy + 1], f[b + g.sxy + g.nx], f[b + g.sxy + g.nx + 1], c[0],
                           c[1], c[2]);
}

// FluidMedium::sample_density_and_temp (src/fluid.rs:739-788) at npts points,
// evaluated in FP64 from a snapshot of density and temperature; out holds
// (density, temperature) pairs.  Aabb::extent = max(max - min, 0) per axis
// (src/math.rs:832-834); Rust's clamp / max as written there.
struct Box3 {
    double lo[3], hi[3];
};

__device__ __forceinline__ double rclamp_d(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_medium_sample(Dims g, Box3 box,
                                                            const T *__restrict__ rho,
                                                            const T *__restrict__ temp,
                                                            const double *__restrict__ xyz,
                                                            long npts, double *__restrict__ out) {
    const long a = (long)blockIdx.x * kThreads + tid_lin();
    if (a >= npts) return;
    const double p[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    if (p[0] < box.lo[0] || p[0] > box.hi[0] || p[1] < box.lo[1] || p[1] > box.hi[1] ||
        p[2] < box.lo[2] || p[2] > box.hi[2]) {
        out[2 * a] = 0.0;
        out[2 * a + 1] = 0.0;
        return;
    }
    double u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ext = fmax(box.hi[d] - box.lo[d], 0.0);
        u[d] = (p[d] - box.lo[d]) / fmax(ext, 1e-9);
    }
    // edge feathering (src/fluid.rs:754-758)
    const double fx = rclamp_d(4.0 * u[0] * (1.0 - u[0]), 0.0, 1.0);
    const double fz = rclamp_d(4.0 * u[2] * (1.0 - u[2]), 0.0, 1.0);
    const double fy = rclamp_d(1.0 - fmax(u[1] - 0.95, 0.0) / 0.05, 0.0, 1.0);
    const double feather = sqrt(fx * fz * fy);
    const int n[3] = {g.nx, g.ny, g.nz};
    int i0[3];
    double c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x = rclamp_d(u[d] * (double)(n[d] - 1), 0.5, (double)n[d] - 1.500001);
        const double fl = floor(x);
        i0[d] = min(max((int)fl, 0), n[d] - 2);  // NaN-safe index
        c[d] = x - fl;
    }
    const int b = (i0[2] * g.ny + i0[1]) * g.nx + i0[0];
    const double dv =
        lerp3<double>(rho[b], rho[b + 1], rho[b + g.nx], rho[b + g.nx + 1], rho[b + g.sxy],
                      rho[b + g.sxy + 1], rho[b + g.sxy + g.nx], rho[b + g.sxy + g.nx + 1],
                      c[0], c[1], c[2]);
    const double tv =
        lerp3<double>(temp[b], temp[b + 1], temp[b + g.nx], temp[b + g.nx + 1], temp[b + g.sxy],
                      temp[b + g.sxy + 1], temp[b + g.sxy + g.nx], temp[b + g.sxy + g.nx + 1],
                      c[0], c[1], c[2]);
    out[2 * a] = dv * feather;
    out[2 * a + 1] = tv * feather;
}

}  // namespace
}  // namespace aether
