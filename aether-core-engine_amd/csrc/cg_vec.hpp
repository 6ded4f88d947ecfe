// cg_vec.hpp -- the two-launch CG iteration (and the red-black SOR pass) with
// four x-cells per thread.
//
// The recurrences of body_cg_dir / body_cg_update (src/fluid.rs:463-487),
// with each a x + y a fused multiply-add, laid out for the memory pipeline:
// every thread owns a 4 (x) x ZU (z) block, moves it with 16-byte loads and
// stores, and takes its x-neighbours from the adjacent lanes (whole-wave DPP
// shifts) instead of from memory; only lanes at a row-segment or wavefront
// edge load their x-neighbour.  The red-black SOR kernels keep the exact
// per-cell arithmetic of k_rbgs.  A
// one-cell-per-thread stencil issues ~10 four-byte memory instructions per
// cell and is limited by the address unit, not by HBM; here it is ~2
// sixteen-byte ones per four cells.  Needs nx % 4 == 0 (rows start
// 16-byte aligned).
#pragma once
#include "fluid_kernels.hpp"

namespace aether {
namespace {

// Workgroups of 256 threads cover txv float4 lanes of ty rows (txv = nx/4
// capped at 64, ty = 256 / txv); threads past txv * ty idle.
struct VTiling {
    int gx, gy, gz, nb, remap;
    int txv, ty;
};

__device__ __forceinline__ void vtile_coords(const VTiling &t, int &bx, int &by, int &bz) {
    int b = blockIdx.x;
    if (t.remap) b = (b & 7) * (t.nb >> 3) + (b >> 3);
    bx = b % t.gx;
    const int rest = b / t.gx;
    by = rest % t.gy;
    bz = rest / t.gy;
}

struct VCol {
    int i0, j, k0;       // first cell (x multiple of 4), row, first plane
    int off;             // in-plane offset of the (clamped) first cell
    int offym, offyp;    // the same for the rows below / above (clamped)
    int lane;
    bool valid;          // owns four in-grid cells
    bool needL, needR;   // x-neighbour comes from memory, not from the adjacent lane
};

template <int ZU>
__device__ __forceinline__ VCol vcol_at(const Dims &g, const VTiling &vt, int bx, int by,
                                        int bz) {
    const int t = threadIdx.x;
    const int txv = t % vt.txv, ty = t / vt.txv;
    VCol c;
    c.i0 = 4 * (bx * vt.txv + txv);
    c.j = by * vt.ty + ty;
    c.k0 = bz * ZU;
    c.valid = ty < vt.ty && c.i0 < g.nx && c.j < g.ny;
    const int ic0 = min(c.i0, g.nx - 4), jc = min(c.j, g.ny - 1);
    c.off = jc * g.nx + ic0;
    c.offym = max(jc - 1, 0) * g.nx + ic0;
    c.offyp = min(jc + 1, g.ny - 1) * g.nx + ic0;
    c.lane = t & 63;
    // lane - 1 holds x-4..x-1 when it is in the same row segment and wave;
    // lane + 1 likewise (an idle or other-row lane only ever sits next to a
    // segment edge, which loads instead)
    c.needL = c.valid && c.i0 > 0 && (c.lane == 0 || txv == 0);
    c.needR = c.valid && c.i0 + 4 < g.nx && (c.lane == 63 || txv == vt.txv - 1);
    return c;
}

// A tiling is regular when one workgroup row spans the whole x extent
// (gx == 1), a wavefront holds whole rows (txv a power of two dividing 64,
// txv * ty == 256), and the rows and planes divide evenly (ny % ty == 0,
// nz % ZU == 0).  Then every thread owns four in-grid cells on every plane,
// and every x-neighbour inside the row is in the adjacent lane of the same
// wavefront (the row ends are walls, which the stencil masks): no edge
// loads, no validity tests, and the coordinates come from shifts.
inline bool vtiling_regular(const VTiling &t, int ny, int nz, int zu) {
    return t.gx == 1 && t.txv > 0 && (t.txv & (t.txv - 1)) == 0 && 64 % t.txv == 0 &&
           t.txv * t.ty == 256 && ny % t.ty == 0 && nz % zu == 0;
}

// the tile (by, bz) of a regular tiling (its bx is 0)
template <int ZU>
__device__ __forceinline__ VCol vcol_regular_at(const Dims &g, const VTiling &vt, int by,
                                                int bz) {
    const int t = threadIdx.x, sh = __builtin_ctz(vt.txv);
    VCol c;
    c.i0 = 4 * (t & (vt.txv - 1));
    c.j = by * vt.ty + (t >> sh);
    c.k0 = bz * ZU;
    c.valid = true;
    c.off = c.j * g.nx + c.i0;
    c.offym = max(c.j - 1, 0) * g.nx + c.i0;
    c.offyp = min(c.j + 1, g.ny - 1) * g.nx + c.i0;
    c.lane = t & 63;
    c.needL = c.needR = false;
    return c;
}

// the tile of this workgroup (renumbered per XCD slab as vtile_coords)
template <int ZU>
__device__ __forceinline__ VCol vcol_regular(const Dims &g, const VTiling &vt) {
    int b = blockIdx.x;
    if (vt.remap) b = (b & 7) * (vt.nb >> 3) + (b >> 3);
    return vcol_regular_at<ZU>(g, vt, b % vt.gy, b / vt.gy);
}

// Output array of a four-wide kernel under one store policy.  ST 0: plain
// stores; 1: non-temporal; 2: write-through, through a buffer resource with
// the sc1 cache-policy bit, so that the stores leave no dirty lines in L2 for
// the end-of-kernel write-back.  Measured on MI355X at 128^3: write-through
// takes the one-launch CG iteration from 28.2 to 22.1 us per launch pair;
// non-temporal stores are slower inside the step (they also bypass the
// Infinity Cache that the next launch reads from).  ST 2 needs the array
// under 2 GiB (32-bit byte offsets); the buffer's range check drops any
// store past nbytes.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
constexpr int kStPlain = 0, kStNT = 1, kStWT = 2;
template <int ST>
struct Out4 {
    float4 *p;
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ Out4(float *ptr, long nbytes) : p(reinterpret_cast<float4 *>(ptr)) {
        if constexpr (ST == kStWT)
            rs = __builtin_amdgcn_make_buffer_rsrc(ptr, (short)0, (int)nbytes, 0x00020000);
    }
    __device__ __forceinline__ void put(int id4, const float4 &v) const {
        if constexpr (ST == kStWT) {
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), rs, id4 * 16, 0,
                                                   16 /* sc1 */);
        } else if constexpr (ST == kStNT) {
            __builtin_nontemporal_store(v.x, &p[id4].x);
            __builtin_nontemporal_store(v.y, &p[id4].y);
            __builtin_nontemporal_store(v.z, &p[id4].z);
            __builtin_nontemporal_store(v.w, &p[id4].w);
        } else {
            p[id4] = v;
        }
    }
};

// Loads of a four-wide kernel under one store policy.  With write-through
// stores (arrays under 2 GiB) the byte offset fits 32 bits, so the address is
// the array base (scalar registers) plus a zero-extended 32-bit offset: one
// shift per load instead of 64-bit address arithmetic.
template <int ST>
__device__ __forceinline__ float4 ld4(const float *__restrict__ base, int o4) {
    if constexpr (ST == kStWT)
        return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) +
                                                 (size_t)((unsigned)o4 << 4));
    else
        return reinterpret_cast<const float4 *>(base)[o4];
}
template <int ST>
__device__ __forceinline__ float ld1(const float *__restrict__ base, int o) {
    if constexpr (ST == kStWT)
        return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) +
                                                (size_t)((unsigned)o << 2));
    else
        return base[o];
}

// a x + y per component, each one fused multiply-add
__device__ __forceinline__ float4 fma4(float a, const float4 &x, const float4 &y) {
    return make_float4(__builtin_fmaf(a, x.x, y.x), __builtin_fmaf(a, x.y, y.y),
                       __builtin_fmaf(a, x.z, y.z), __builtin_fmaf(a, x.w, y.w));
}

__device__ __forceinline__ float f4(const float4 &v, int e) {
    return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}
__device__ __forceinline__ float4 sub4(const float4 &x, const float4 &y) {
    return make_float4(x.x - y.x, x.y - y.y, x.z - y.z, x.w - y.w);
}
// value of lane ^ 32 (the other half of the wavefront) by v_permlane32_swap:
// its first result holds lanes 0-31 in both halves, its second lanes 32-63
__device__ __forceinline__ float other_half(float v, bool hi) {
    const unsigned b = __float_as_uint(v);
    const auto r = __builtin_amdgcn_permlane32_swap(b, b, false, false);
    return __uint_as_float(hi ? r[0] : r[1]);
}
// value of lane - 1 and of lane + 1 through a whole-wavefront DPP shift (no
// LDS round trip); lane 0 and lane 63 keep their own value
__device__ __forceinline__ float lane_prev(float v) {
    const int b = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_next(float v) {
    const int b = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

// A x on the four cells of plane k of a column (lap7 per cell, interior only).
// Every cell's stencil is formed and the non-interior ones are then zeroed by
// a select, not skipped by a branch: the stencil reads registers only, and a
// branch per cell would also pull the neighbours' arithmetic into it.
__device__ __forceinline__ void lap4(const Dims &g, const VCol &c, int k, const float4 &x,
                                     float xl, float xr, const float4 &ym, const float4 &yp,
                                     const float4 &zm, const float4 &zp, float (&q)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = c.i0 + e;
        const float v =
            lap7<float>(f4(x, e), e == 0 ? xl : f4(x, e - 1), e == 3 ? xr : f4(x, e + 1),
                        f4(ym, e), f4(yp, e), f4(zm, e), f4(zp, e), i > 1, i + 2 < g.nx, c.j > 1,
                        c.j + 2 < g.ny, k > 1, k + 2 < g.nz);
        q[e] = is_interior(g, i, c.j, k) ? v : 0.0f;
    }
}

// CG launch 1 of iteration `it` (as body_cg_dir): stop rules on r.r, then
// d = r + beta d_old, q = A d, partial d.q; writes d.  NQ: partial sums per
// thread held in registers (256 NQ >= nb keeps their reload one round trip).
// `slot` is the workgroup's partial-sum slot (0 leads), nb_in the number of
// partials the previous launch wrote.
template <int ZU, bool FIRST, int NQ, int ST, bool PEND>
__device__ __forceinline__ void body_cg_dir4(
    const VCol &c, int slot, Dims g, int it, double rel_tol2, int nb_in,
    const double *__restrict__ part_rr, const double *__restrict__ part_bb, CgState *st,
    const float *__restrict__ r, const float *__restrict__ dold, float *__restrict__ dnew,
    double *__restrict__ part_dq, int *done_mirror, int solve_id, float *__restrict__ p) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> prr, pbb;
    prr.load(part_rr, nb_in);
    if (FIRST) pbb.load(part_bb, nb_in);
    const double delta_prev = FIRST ? 0.0 : st->delta[(it - 1) & 1];
    const double b2_prev = FIRST ? 0.0 : st->b2;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 rz[ZU + 2], dz[ZU + 2], ry[ZU][2], dy[ZU][2];
    float xr[ZU][2], xd[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy;
        rz[a] = ld4<ST>(r, (base + c.off) >> 2);
        dz[a] = FIRST ? z4 : ld4<ST>(dold, (base + c.off) >> 2);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        ry[a][0] = ld4<ST>(r, (base + c.offym) >> 2);
        ry[a][1] = ld4<ST>(r, (base + c.offyp) >> 2);
        dy[a][0] = FIRST ? z4 : ld4<ST>(dold, (base + c.offym) >> 2);
        dy[a][1] = FIRST ? z4 : ld4<ST>(dold, (base + c.offyp) >> 2);
        xr[a][0] = xr[a][1] = xd[a][0] = xd[a][1] = 0.0f;
        if (c.needL) {
            xr[a][0] = ld1<ST>(r, base + c.off - 1);
            if (!FIRST) xd[a][0] = ld1<ST>(dold, base + c.off - 1);
        }
        if (c.needR) {
            xr[a][1] = ld1<ST>(r, base + c.off + 4);
            if (!FIRST) xd[a][1] = ld1<ST>(dold, base + c.off + 4);
        }
    }
    // stopping rules (every workgroup reaches the same decision); the stop
    // word holds the stopping launch + 1 (launch 2 it here), so only later
    // launches take the early exit
    const bool lead = slot == 0 && threadIdx.x == 0;
    double beta = 0.0;
    if (threadIdx.x == 0) flag_lds[0] = flag_reg != 0 && flag_reg <= 2 * it;
    const double rr = prr.total(part_rr, nb_in, lds);  // its barrier publishes flag_lds
    if (flag_lds[0]) return;
    if (FIRST) {
        const double bb = pbb.total(part_bb, nb_in, lds);
        const bool stop = !(rr > 1e-24) || (rel_tol2 > 0.0 && rr <= rel_tol2 * bb);
        if (lead) {
            st->b2 = bb;
            st->r2_init = rr;
            st->delta[0] = rr;
            if (stop) {
                st->done = 2 * it + 1;
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
                st->done = 2 * it + 1;
                st->iters = it;
                st->r2_final = rr;
                st->converged = 1;
                *done_mirror = solve_id;
            } else {
                st->delta[it & 1] = rr;
            }
        }
        if (stop) {
            if (PEND) {  // p += alpha_(it-1) d_(it-1), left pending by update launch it - 1
                const float tap = (float)st->cg_alpha[(it - 1) & 1];
                float4 *p4 = reinterpret_cast<float4 *>(p);
#pragma unroll
                for (int a = 0; a < ZU; ++a) {
                    const int k = c.k0 + a;
                    if (c.valid && k < g.nz) {
                        const int id4 = (k * g.sxy + c.off) >> 2;
                        p4[id4] = fma4(tap, dz[a + 1], p4[id4]);
                    }
                }
            }
            return;
        }
        beta = rr / delta_prev;
    }
    const float tb = (float)beta;
    float4 dc[ZU + 2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) dc[a] = FIRST ? rz[a] : fma4(tb, dz[a], rz[a]);
    double acc[1] = {0.0};
    const Out4<ST> o_d(dnew, (long)g.sxy * g.nz * 4);
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const float4 ym = FIRST ? ry[a][0] : fma4(tb, dy[a][0], ry[a][0]);
        const float4 yp = FIRST ? ry[a][1] : fma4(tb, dy[a][1], ry[a][1]);
        const float sl = lane_prev(dc[a + 1].w), sr = lane_next(dc[a + 1].x);
        const float xl = c.needL ? (FIRST ? xr[a][0] : __builtin_fmaf(tb, xd[a][0], xr[a][0])) : sl;
        const float xrv = c.needR ? (FIRST ? xr[a][1] : __builtin_fmaf(tb, xd[a][1], xr[a][1])) : sr;
        if (c.valid && k < g.nz) {
            float q[4];
            lap4(g, c, k, dc[a + 1], xl, xrv, ym, yp, dc[a], dc[a + 2], q);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                acc[0] = __builtin_fma((double)f4(dc[a + 1], e), (double)q[e], acc[0]);
            o_d.put((k * g.sxy + c.off) >> 2, dc[a + 1]);
        }
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) part_dq[slot] = acc[0];
}

template <int ZU, bool FIRST, int NQ, int ST, bool PEND>
__global__ __launch_bounds__(kThreads) void k_cg_dir4(
    Dims g, VTiling vt, int it, double rel_tol2, int nb_in, const double *__restrict__ part_rr,
    const double *__restrict__ part_bb, CgState *st, const float *__restrict__ r,
    const float *__restrict__ dold, float *__restrict__ dnew, double *__restrict__ part_dq,
    int *done_mirror, int solve_id, float *__restrict__ p) {
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    body_cg_dir4<ZU, FIRST, NQ, ST, PEND>(vcol_at<ZU>(g, vt, bx, by, bz), blockIdx.x, g, it,
                                          rel_tol2, nb_in, part_rr, part_bb, st, r, dold, dnew,
                                          part_dq, done_mirror, solve_id, p);
}

// CG launch 2 of iteration `it` (as body_cg_update): alpha from d.q,
// p += alpha d, r -= alpha A d (A d recomputed), partial r.r.
// PM: how p is brought up to date.  kPEvery: p += alpha d every launch.
// kPSkip (even it): not at all (alpha_it d_it stays pending).  kPBoth (odd
// it): p <- (p + alpha_(it-1) d_(it-1)) + alpha_it d_it, with d_(it-1) in
// dprev (the other d buffer): the same two FP32 additions in the same order
// as kPEvery, so every mode gives bit-identical p.  The lead stores alpha
// for the launches and the tail that apply a pending term.
constexpr int kPEvery = 0, kPSkip = 1, kPBoth = 2;
template <int ZU, int NQ, int ST, int PM>
__device__ __forceinline__ void body_cg_update4(
    const VCol &c, int slot, int nb, Dims g, int it, const double *__restrict__ part_dq,
    CgState *st, const float *__restrict__ d, float *__restrict__ p, float *__restrict__ r,
    double *__restrict__ part_rr, int *done_mirror, int solve_id,
    const float *__restrict__ dprev) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> pdq;
    pdq.load(part_dq, nb);
    const double delta = st->delta[it & 1];
    const Out4<ST> o_p(p, (long)g.sxy * g.nz * 4), o_r(r, (long)g.sxy * g.nz * 4);
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 dz[ZU + 2], dy[ZU][2], pc[ZU], rc[ZU], dp[ZU];
    float xd[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy;
        dz[a] = ld4<ST>(d, (base + c.off) >> 2);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        dy[a][0] = ld4<ST>(d, (base + c.offym) >> 2);
        dy[a][1] = ld4<ST>(d, (base + c.offyp) >> 2);
        pc[a] = PM == kPSkip ? z4 : ld4<ST>(p, (base + c.off) >> 2);
        dp[a] = PM == kPBoth ? ld4<ST>(dprev, (base + c.off) >> 2) : z4;
        rc[a] = ld4<ST>(r, (base + c.off) >> 2);
        xd[a][0] = xd[a][1] = 0.0f;
        if (c.needL) xd[a][0] = ld1<ST>(d, base + c.off - 1);
        if (c.needR) xd[a][1] = ld1<ST>(d, base + c.off + 4);
    }
    const bool lead = slot == 0 && threadIdx.x == 0;
    // the stop word holds the stopping launch + 1 (launch 2 it + 1 here)
    if (threadIdx.x == 0) flag_lds[0] = flag_reg != 0 && flag_reg <= 2 * it + 1;
    const double dq = pdq.total(part_dq, nb, lds);
    if (flag_lds[0]) return;
    const float tap = PM == kPBoth ? (float)st->cg_alpha[(it - 1) & 1] : 0.0f;
    if (fabs(dq) < 1e-28) {
        if (lead) {
            st->done = 2 * it + 2;
            st->iters = it;
            st->converged = 1;
            st->r2_final = delta;
            *done_mirror = solve_id;
        }
        if (PM == kPBoth) {  // the pending term of launch it - 1
#pragma unroll
            for (int a = 0; a < ZU; ++a) {
                const int k = c.k0 + a;
                if (c.valid && k < g.nz) o_p.put((k * g.sxy + c.off) >> 2, fma4(tap, dp[a], pc[a]));
            }
        }
        return;
    }
    const double alpha = delta / dq;
    if (lead) st->cg_alpha[it & 1] = alpha;
    const float ta = (float)alpha;
    double acc[1] = {0.0};
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const float sl = lane_prev(dz[a + 1].w), sr = lane_next(dz[a + 1].x);
        const float xl = c.needL ? xd[a][0] : sl;
        const float xrv = c.needR ? xd[a][1] : sr;
        if (c.valid && k < g.nz) {
            float q[4];
            lap4(g, c, k, dz[a + 1], xl, xrv, dy[a][0], dy[a][1], dz[a], dz[a + 2], q);
            const float4 pn = PM == kPBoth ? fma4(ta, dz[a + 1], fma4(tap, dp[a], pc[a]))
                                           : fma4(ta, dz[a + 1], pc[a]);
            const float4 rn = fma4(-ta, make_float4(q[0], q[1], q[2], q[3]), rc[a]);
            acc[0] = __builtin_fma((double)rn.x, (double)rn.x, acc[0]);
            acc[0] = __builtin_fma((double)rn.y, (double)rn.y, acc[0]);
            acc[0] = __builtin_fma((double)rn.z, (double)rn.z, acc[0]);
            acc[0] = __builtin_fma((double)rn.w, (double)rn.w, acc[0]);
            const int id4 = (k * g.sxy + c.off) >> 2;
            if (PM != kPSkip) o_p.put(id4, pn);
            o_r.put(id4, rn);
        }
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) part_rr[slot] = acc[0];
}

// nb_in: the number of d.q partials to sum (vt.nb, or 1 when a separate
// kernel has already reduced them to one total).
template <int ZU, int NQ, int ST, int PM>
__global__ __launch_bounds__(kThreads) void k_cg_update4(
    Dims g, VTiling vt, int it, int nb_in, const double *__restrict__ part_dq, CgState *st,
    const float *__restrict__ d, float *__restrict__ p, float *__restrict__ r,
    double *__restrict__ part_rr, int *done_mirror, int solve_id,
    const float *__restrict__ dprev) {
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    body_cg_update4<ZU, NQ, ST, PM>(vcol_at<ZU>(g, vt, bx, by, bz), blockIdx.x, nb_in, g, it,
                                    part_dq, st, d, p, r, part_rr, done_mirror, solve_id, dprev);
}

// Closes a four-wide solve whose iteration cap ended it after an even launch
// (that launch left alpha_(it) d_(it) pending): p += alpha d.  Skipped when a
// launch stopped the solve (that launch applied the pending term).
__global__ __launch_bounds__(kThreads) void k_cg_p_tail(const CgState *st, int last_it,
                                                        const float4 *__restrict__ d,
                                                        float4 *__restrict__ p, long n4) {
    if (st->done) return;
    const float ta = (float)st->cg_alpha[last_it & 1];
    for (long a = (long)blockIdx.x * kThreads + threadIdx.x; a < n4; a += (long)gridDim.x * kThreads)
        p[a] = fma4(ta, d[a], p[a]);
}

// Red-black SOR update of the colour-`colour` interior cells among four
// x-cells (src/fluid.rs:400-437; per-cell arithmetic as k_rbgs): ctr holds
// the cells, left/right their x-neighbours outside the float4, ym/yp/zm/zp
// the rows and planes beside them, bb the right-hand side.  Cells of the
// other colour or off the interior keep their bits.
__device__ __forceinline__ float4 sor_colour4(const Dims &g, int i0, int j, int k, int colour,
                                              float omega, const float4 &ctr, float left,
                                              float right, const float4 &ym, const float4 &yp,
                                              const float4 &zm, const float4 &zp,
                                              const float4 &bb, bool &any) {
    float o[4] = {ctr.x, ctr.y, ctr.z, ctr.w};
    any = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = i0 + e;
        if (!is_interior(g, i, j, k) || ((i + j + k) & 1) != colour) continue;
        any = true;
        float diag = 0.0f, nb = 0.0f;
        if (i > 1) { diag += 1.0f; nb += e == 0 ? left : f4(ctr, e - 1); }
        if (i + 2 < g.nx) { diag += 1.0f; nb += e == 3 ? right : f4(ctr, e + 1); }
        if (j > 1) { diag += 1.0f; nb += f4(ym, e); }
        if (j + 2 < g.ny) { diag += 1.0f; nb += f4(yp, e); }
        if (k > 1) { diag += 1.0f; nb += f4(zm, e); }
        if (k + 2 < g.nz) { diag += 1.0f; nb += f4(zp, e); }
        const float gs = sor_quot(f4(bb, e) + nb, diag);
        o[e] = (1.0f - omega) * o[e] + omega * gs;
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

// Red-black SOR pass of one colour (src/fluid.rs:400-437) on the four-wide
// layout: the same per-cell arithmetic as k_rbgs.  A cell's neighbours all
// have the other colour, which this pass leaves unchanged, so every thread
// reads consistent values and writes back its whole float4 (the other-colour
// cells with their unchanged bits).
template <int ZU, int ST>
__global__ __launch_bounds__(kThreads) void k_rbgs4(Dims g, VTiling vt, int color, float omega,
                                                    const float *__restrict__ b,
                                                    float *__restrict__ p) {
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol_at<ZU>(g, vt, bx, by, bz);
    const Out4<ST> o_p(p, (long)g.sxy * g.nz * 4);
    float4 pz[ZU + 2], py[ZU][2], bc[ZU];
    float xl[ZU], xr[ZU];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy;
        pz[a] = ld4<ST>(p, (base + c.off) >> 2);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        py[a][0] = ld4<ST>(p, (base + c.offym) >> 2);
        py[a][1] = ld4<ST>(p, (base + c.offyp) >> 2);
        bc[a] = ld4<ST>(b, (base + c.off) >> 2);
        xl[a] = xr[a] = 0.0f;
        if (c.needL) xl[a] = p[base + c.off - 1];
        if (c.needR) xr[a] = p[base + c.off + 4];
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const float sl = lane_prev(pz[a + 1].w), sr = lane_next(pz[a + 1].x);
        const float left = c.needL ? xl[a] : sl, right = c.needR ? xr[a] : sr;
        if (c.valid && k < g.nz) {
            bool any;
            const float4 o = sor_colour4(g, c.i0, c.j, k, color, omega, pz[a + 1], left, right,
                                         py[a][0], py[a][1], pz[a], pz[a + 2], bc[a], any);
            if (any) o_p.put((k * g.sxy + c.off) >> 2, o);
        }
    }
}

// One whole red-black SOR sweep (src/fluid.rs:400-437) in one launch, from
// p_in to p_out.  Each thread forms, from p_in, the colour-0 values of its
// own column and one plane past each end of it, and shares its own planes
// through LDS; the threads on the first and last row of the workgroup's
// tile also form the colour-0 values of the row beyond the tile.  Each
// thread then updates its colour-1 cells and writes its whole column.
// Those colour-0 values are what the colour-0 pass of the two-launch sweep
// would store there (same inputs, same arithmetic), so the sweep is
// bit-identical to two k_rbgs4 passes.  Neighbouring workgroups form the
// rows and planes between them twice, which is why the sweep reads p_in and
// writes a second array.  Needs every grid row to lie in one wave with no
// lane past its end (nx = 4 w4, w4 dividing 64): then x-neighbours always
// come from adjacent lanes and no edge loads are needed.
template <int ZU, int ST>
__global__ __launch_bounds__(kThreads) void k_rbsor4(Dims g, VTiling vt, float omega,
                                                     const float *__restrict__ b,
                                                     const float *__restrict__ p_in,
                                                     float *__restrict__ p_out) {
    __shared__ float4 lds_rc[ZU][kThreads];
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol_at<ZU>(g, vt, bx, by, bz);
    const Out4<ST> o_p(p_out, (long)g.sxy * g.nz * 4);
    const int t = threadIdx.x, ty = t / vt.txv;
    const bool first_row = ty == 0, last_row = ty == vt.ty - 1;
    const int jc = min(c.j, g.ny - 1), ic0 = c.off - jc * g.nx;
    auto row = [&](int dj) { return min(max(c.j + dj, 0), g.ny - 1) * g.nx + ic0; };
    auto plane = [&](int k) { return min(max(k, 0), g.nz - 1) * g.sxy; };
    float4 C[ZU + 4];      // row j, planes k0-2 .. k0+ZU+1
    float4 Y1[2][ZU + 2];  // rows j-1, j+1, planes k0-1 .. k0+ZU
    float4 BC[ZU + 2];     // b on row j, planes k0-1 .. k0+ZU
#pragma unroll
    for (int a = 0; a < ZU + 4; ++a) C[a] = ld4<ST>(p_in, (plane(c.k0 - 2 + a) + c.off) >> 2);
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = plane(c.k0 - 1 + a);
        Y1[0][a] = ld4<ST>(p_in, (base + c.offym) >> 2);
        Y1[1][a] = ld4<ST>(p_in, (base + c.offyp) >> 2);
        BC[a] = ld4<ST>(b, (base + c.off) >> 2);
    }
    // the first and last row of the tile also need p two rows away and b on
    // the row beyond the tile; loaded with the rest, so that the colour-0
    // values of that row cost no second memory round trip
    const int dj_edge = first_row ? -1 : last_row ? 1 : 0;
    float4 FAR[ZU], BB[ZU];
    if (dj_edge != 0) {
        const int rfar = row(2 * dj_edge), rnear = dj_edge < 0 ? c.offym : c.offyp;
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            const int base = plane(c.k0 + a);
            FAR[a] = ld4<ST>(p_in, (base + rfar) >> 2);
            BB[a] = ld4<ST>(b, (base + rnear) >> 2);
        }
    }
    bool any;
    // colour 0 on row j, planes k0-1 .. k0+ZU; the thread's own planes to LDS.
    // x-neighbours from the adjacent lanes by whole-wave shifts: the launch
    // condition puts whole rows in each wavefront (row ends are walls)
    float4 RC[ZU + 2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const float sl = lane_prev(C[a + 1].w), sr = lane_next(C[a + 1].x);
        RC[a] = sor_colour4(g, c.i0, c.j, c.k0 - 1 + a, 0, omega, C[a + 1], sl, sr, Y1[0][a],
                            Y1[1][a], C[a], C[a + 2], BC[a], any);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) lds_rc[a][t] = RC[a + 1];
    // x-neighbours on rows j-1 and j+1 (every lane takes part in the exchange)
    float yl[2][ZU], yr[2][ZU];
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            yl[s2][a] = lane_prev(Y1[s2][a + 1].w);
            yr[s2][a] = lane_next(Y1[s2][a + 1].x);
        }
    // colour 0 on the row beyond the tile (first and last tile rows only)
    float4 RE[ZU];
#pragma unroll
    for (int a = 0; a < ZU; ++a) RE[a] = make_float4(0.f, 0.f, 0.f, 0.f);
    // (one branch per side, each with constant array indices, so that the
    // arrays stay in registers)
    auto edge_row = [&](const float4(&y)[ZU + 2], const float(&l)[ZU], const float(&r)[ZU],
                        int dj) {
#pragma unroll
        for (int a = 0; a < ZU; ++a)
            RE[a] = sor_colour4(g, c.i0, c.j + dj, c.k0 + a, 0, omega, y[a + 1], l[a], r[a],
                                dj < 0 ? FAR[a] : C[a + 2], dj < 0 ? C[a + 2] : FAR[a], y[a],
                                y[a + 2], BB[a], any);
    };
    if (first_row) edge_row(Y1[0], yl[0], yr[0], -1);
    if (last_row) edge_row(Y1[1], yl[1], yr[1], 1);
    __syncthreads();
    // colour 1 on the thread's own cells, from the colour-0 values above
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const float sl = lane_prev(RC[a + 1].w), sr = lane_next(RC[a + 1].x);
        // (indices kept in range on the rows that take RE instead)
        const float4 ym = first_row ? RE[a] : lds_rc[a][max(t - vt.txv, 0)];
        const float4 yp = last_row ? RE[a] : lds_rc[a][min(t + vt.txv, kThreads - 1)];
        if (c.valid && k < g.nz) {
            const float4 o = sor_colour4(g, c.i0, c.j, k, 1, omega, RC[a + 1], sl, sr, ym, yp,
                                         RC[a], RC[a + 2], BC[a + 1], any);
            o_p.put((k * g.sxy + c.off) >> 2, o);  // every cell: p_out holds the whole sweep
        }
    }
}

}  // namespace
}  // namespace aether
