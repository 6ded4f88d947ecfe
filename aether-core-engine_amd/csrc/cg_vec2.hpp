// cg_vec2.hpp -- the two-launch CG iteration in FP64 with two x-cells per
// thread.
//
// The recurrences of body_cg_dir / body_cg_update (src/fluid.rs:463-487)
// with the FP64 instantiation's per-cell arithmetic unchanged: every update
// is a product then a sum (the library builds with -ffp-contract=off), the
// stencil is lap7<double> in the reference's order, and the dot products
// add each cell's product in FP64.  Only the layout differs from the
// one-cell-per-thread kernels: every thread owns a 2 (x) x ZU (z) block and
// moves it with 16-byte double2 loads and stores (the FP32 four-wide form's
// memory pattern), and takes its x-neighbours from the adjacent lanes by
// whole-wave DPP shifts of both halves of a double; lanes at a row-segment
// or wavefront edge load theirs.  The partial sums follow this tiling, so
// the dot products add the same terms in another order than the
// one-cell-per-thread form: the iterates agree to FP64 rounding.  Needs
// nx % 2 == 0 (rows start 16-byte aligned).
#pragma once
#include "cg_vec.hpp"

namespace aether {
namespace {

// The tile column of this thread in a two-wide tiling: txv double2 lanes of
// ty rows per workgroup (vt as VTiling, with two cells per lane).
template <int ZU>
__device__ __forceinline__ VCol vcol2_at(const Dims &g, const VTiling &vt, int bx, int by,
                                         int bz) {
    const int t = threadIdx.x;
    const int txv = t % vt.txv, ty = t / vt.txv;
    VCol c;
    c.i0 = 2 * (bx * vt.txv + txv);
    c.j = by * vt.ty + ty;
    c.k0 = bz * ZU;
    c.valid = ty < vt.ty && c.i0 < g.nx && c.j < g.ny;
    const int ic0 = min(c.i0, g.nx - 2), jc = min(c.j, g.ny - 1);
    c.off = jc * g.nx + ic0;
    c.offym = max(jc - 1, 0) * g.nx + ic0;
    c.offyp = min(jc + 1, g.ny - 1) * g.nx + ic0;
    c.lane = t & 63;
    // lane - 1 holds x-2..x-1 when it is in the same row segment and wave
    c.needL = c.valid && c.i0 > 0 && (c.lane == 0 || txv == 0);
    c.needR = c.valid && c.i0 + 2 < g.nx && (c.lane == 63 || txv == vt.txv - 1);
    return c;
}

// Loads and stores of a two-wide kernel.  With write-through stores (arrays
// under 2 GiB) the byte offset fits 32 bits, as in ld4 and Out4.
template <int ST>
__device__ __forceinline__ double2 ld2(const double *__restrict__ base, int o2) {
    if constexpr (ST == kStWT)
        return *reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(base) +
                                                  (size_t)((unsigned)o2 << 4));
    else
        return reinterpret_cast<const double2 *>(base)[o2];
}
template <int ST>
__device__ __forceinline__ double ld1d(const double *__restrict__ base, int o) {
    if constexpr (ST == kStWT)
        return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) +
                                                 (size_t)((unsigned)o << 3));
    else
        return base[o];
}
template <int ST>
struct Out2 {
    double2 *p;
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ Out2(double *ptr, long nbytes)
        : p(reinterpret_cast<double2 *>(ptr)) {
        if constexpr (ST == kStWT)
            rs = __builtin_amdgcn_make_buffer_rsrc(ptr, (short)0, (int)nbytes, 0x00020000);
    }
    __device__ __forceinline__ void put(int id2, const double2 &v) const {
        if constexpr (ST == kStWT)
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), rs, id2 * 16, 0,
                                                   16 /* sc1 */);
        else
            p[id2] = v;
    }
};

// a x + y per component: a product, then a sum (as the scalar FP64 kernels)
__device__ __forceinline__ double2 axpy2(double a, const double2 &x, const double2 &y) {
    const double px = a * x.x, py = a * x.y;
    return make_double2(px + y.x, py + y.y);
}

// value of lane - 1 / lane + 1 by whole-wavefront DPP shifts of both 32-bit
// halves; lane 0 / lane 63 keep their own value
__device__ __forceinline__ double lane_prev_d(double v) {
    const unsigned long long b = __double_as_longlong(v);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const unsigned l = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, 0x138 /* wave_shr:1 */, 0xf,
                                                             0xf, false);
    const unsigned h = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
}
__device__ __forceinline__ double lane_next_d(double v) {
    const unsigned long long b = __double_as_longlong(v);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const unsigned l = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, 0x130 /* wave_shl:1 */, 0xf,
                                                             0xf, false);
    const unsigned h = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
}

// A x on the two cells of plane k of a column: lap7 per interior cell, with
// the neighbours in the scalar kernels' order (x-, x+, y-, y+, z-, z+); the
// other cells give 0.
__device__ __forceinline__ void lap2(const Dims &g, const VCol &c, int k, const double2 &x,
                                     double xl, double xr, const double2 &ym, const double2 &yp,
                                     const double2 &zm, const double2 &zp, double (&q)[2]) {
    const double xc[2] = {x.x, x.y}, xm[2] = {xl, x.x}, xp[2] = {x.y, xr};
    const double vym[2] = {ym.x, ym.y}, vyp[2] = {yp.x, yp.y};
    const double vzm[2] = {zm.x, zm.y}, vzp[2] = {zp.x, zp.y};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int i = c.i0 + e;
        const double v = lap7<double>(xc[e], xm[e], xp[e], vym[e], vyp[e], vzm[e], vzp[e], i > 1,
                                      i + 2 < g.nx, c.j > 1, c.j + 2 < g.ny, k > 1, k + 2 < g.nz);
        q[e] = is_interior(g, i, c.j, k) ? v : 0.0;
    }
}

// CG launch 1 of iteration `it` (as body_cg_dir and body_cg_dir4): stop
// rules on r.r, d = r + beta d_old, q = A d, partial d.q; writes d.  PEND:
// p += alpha_(it-1) d_(it-1) is pending (the update launch of an even
// iteration skips p) and a stop here applies it.
template <int ZU, bool FIRST, int NQ, int ST, bool PEND>
__global__ __launch_bounds__(kThreads) void k_cg_dir2(
    Dims g, VTiling vt, int it, double rel_tol2, int nb_in, const double *__restrict__ part_rr,
    const double *__restrict__ part_bb, CgState *st, const double *__restrict__ r,
    const double *__restrict__ dold, double *__restrict__ dnew, double *__restrict__ part_dq,
    int *done_mirror, int solve_id, double *__restrict__ p) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol2_at<ZU>(g, vt, bx, by, bz);
    const int slot = blockIdx.x;
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> prr, pbb;
    prr.load(part_rr, nb_in);
    if (FIRST) pbb.load(part_bb, nb_in);
    const double delta_prev = FIRST ? 0.0 : st->delta[(it - 1) & 1];
    const double b2_prev = FIRST ? 0.0 : st->b2;
    const double2 z2 = make_double2(0.0, 0.0);
    double2 rz[ZU + 2], dz[ZU + 2], ry[ZU][2], dy[ZU][2];
    double xr[ZU][2], xd[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy;
        rz[a] = ld2<ST>(r, (base + c.off) >> 1);
        dz[a] = FIRST ? z2 : ld2<ST>(dold, (base + c.off) >> 1);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        ry[a][0] = ld2<ST>(r, (base + c.offym) >> 1);
        ry[a][1] = ld2<ST>(r, (base + c.offyp) >> 1);
        dy[a][0] = FIRST ? z2 : ld2<ST>(dold, (base + c.offym) >> 1);
        dy[a][1] = FIRST ? z2 : ld2<ST>(dold, (base + c.offyp) >> 1);
        xr[a][0] = xr[a][1] = xd[a][0] = xd[a][1] = 0.0;
        if (c.needL) {
            xr[a][0] = ld1d<ST>(r, base + c.off - 1);
            if (!FIRST) xd[a][0] = ld1d<ST>(dold, base + c.off - 1);
        }
        if (c.needR) {
            xr[a][1] = ld1d<ST>(r, base + c.off + 2);
            if (!FIRST) xd[a][1] = ld1d<ST>(dold, base + c.off + 2);
        }
    }
    // stopping rules (every workgroup reaches the same decision); the stop
    // word holds the stopping launch + 1 (launch 2 it here)
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
                const double tap = st->cg_alpha[(it - 1) & 1];
                double2 *p2 = reinterpret_cast<double2 *>(p);
#pragma unroll
                for (int a = 0; a < ZU; ++a) {
                    const int k = c.k0 + a;
                    if (c.valid && k < g.nz) {
                        const int id2 = (k * g.sxy + c.off) >> 1;
                        p2[id2] = axpy2(tap, dz[a + 1], p2[id2]);
                    }
                }
            }
            return;
        }
        beta = rr / delta_prev;
    }
    const double tb = beta;
    double2 dc[ZU + 2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) dc[a] = FIRST ? rz[a] : axpy2(tb, dz[a], rz[a]);
    double acc[1] = {0.0};
    const Out2<ST> o_d(dnew, (long)g.sxy * g.nz * 8);
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const double2 ym = FIRST ? ry[a][0] : axpy2(tb, dy[a][0], ry[a][0]);
        const double2 yp = FIRST ? ry[a][1] : axpy2(tb, dy[a][1], ry[a][1]);
        const double sl = lane_prev_d(dc[a + 1].y), sr = lane_next_d(dc[a + 1].x);
        double xl = sl, xrv = sr;
        if (c.needL) xl = FIRST ? xr[a][0] : tb * xd[a][0] + xr[a][0];
        if (c.needR) xrv = FIRST ? xr[a][1] : tb * xd[a][1] + xr[a][1];
        if (c.valid && k < g.nz) {
            double q[2];
            lap2(g, c, k, dc[a + 1], xl, xrv, ym, yp, dc[a], dc[a + 2], q);
            acc[0] += dc[a + 1].x * q[0];
            acc[0] += dc[a + 1].y * q[1];
            o_d.put((k * g.sxy + c.off) >> 1, dc[a + 1]);
        }
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) part_dq[slot] = acc[0];
}

// CG launch 2 of iteration `it` (as body_cg_update and body_cg_update4):
// alpha from d.q, p += alpha d, r -= alpha A d (A d recomputed), partial
// r.r.  PM as body_cg_update4: kPSkip (even it) leaves alpha_it d_it
// pending, kPBoth (odd it) applies p <- (p + alpha_(it-1) d_(it-1)) +
// alpha_it d_it with d_(it-1) in dprev: the same two FP64 sums in the same
// order as updating p every launch.
template <int ZU, int NQ, int ST, int PM>
__global__ __launch_bounds__(kThreads) void k_cg_update2(
    Dims g, VTiling vt, int it, int nb, const double *__restrict__ part_dq, CgState *st,
    const double *__restrict__ d, double *__restrict__ p, double *__restrict__ r,
    double *__restrict__ part_rr, int *done_mirror, int solve_id,
    const double *__restrict__ dprev) {
    __shared__ double lds[8];
    __shared__ int flag_lds[1];
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol2_at<ZU>(g, vt, bx, by, bz);
    const int slot = blockIdx.x;
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> pdq;
    pdq.load(part_dq, nb);
    const double delta = st->delta[it & 1];
    const long nbytes = (long)g.sxy * g.nz * 8;
    const Out2<ST> o_p(p, nbytes), o_r(r, nbytes);
    const double2 z2 = make_double2(0.0, 0.0);
    double2 dz[ZU + 2], dy[ZU][2], pc[ZU], rc[ZU], dp[ZU];
    double xd[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy;
        dz[a] = ld2<ST>(d, (base + c.off) >> 1);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        dy[a][0] = ld2<ST>(d, (base + c.offym) >> 1);
        dy[a][1] = ld2<ST>(d, (base + c.offyp) >> 1);
        pc[a] = PM == kPSkip ? z2 : ld2<ST>(p, (base + c.off) >> 1);
        dp[a] = PM == kPBoth ? ld2<ST>(dprev, (base + c.off) >> 1) : z2;
        rc[a] = ld2<ST>(r, (base + c.off) >> 1);
        xd[a][0] = xd[a][1] = 0.0;
        if (c.needL) xd[a][0] = ld1d<ST>(d, base + c.off - 1);
        if (c.needR) xd[a][1] = ld1d<ST>(d, base + c.off + 2);
    }
    const bool lead = slot == 0 && threadIdx.x == 0;
    // the stop word holds the stopping launch + 1 (launch 2 it + 1 here)
    if (threadIdx.x == 0) flag_lds[0] = flag_reg != 0 && flag_reg <= 2 * it + 1;
    const double dq = pdq.total(part_dq, nb, lds);
    if (flag_lds[0]) return;
    const double tap = PM == kPBoth ? st->cg_alpha[(it - 1) & 1] : 0.0;
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
                if (c.valid && k < g.nz) o_p.put((k * g.sxy + c.off) >> 1, axpy2(tap, dp[a], pc[a]));
            }
        }
        return;
    }
    const double alpha = delta / dq;
    if (lead) st->cg_alpha[it & 1] = alpha;
    const double ta = alpha;
    double acc[1] = {0.0};
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const double sl = lane_prev_d(dz[a + 1].y), sr = lane_next_d(dz[a + 1].x);
        const double xl = c.needL ? xd[a][0] : sl;
        const double xrv = c.needR ? xd[a][1] : sr;
        if (c.valid && k < g.nz) {
            double q[2];
            lap2(g, c, k, dz[a + 1], xl, xrv, dy[a][0], dy[a][1], dz[a], dz[a + 2], q);
            const double2 pn = PM == kPBoth ? axpy2(ta, dz[a + 1], axpy2(tap, dp[a], pc[a]))
                                            : axpy2(ta, dz[a + 1], pc[a]);
            // r - alpha q, as the scalar kernel's rc - ta * q
            const double qa = ta * q[0], qb = ta * q[1];
            const double2 rn = make_double2(rc[a].x - qa, rc[a].y - qb);
            acc[0] += rn.x * rn.x;
            acc[0] += rn.y * rn.y;
            const int id2 = (k * g.sxy + c.off) >> 1;
            if (PM != kPSkip) o_p.put(id2, pn);
            o_r.put(id2, rn);
        }
    }
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) part_rr[slot] = acc[0];
}

// Red-black SOR update of the colour-`colour` interior cells among two
// x-cells (src/fluid.rs:400-437), with k_rbgs's per-cell FP64 arithmetic:
// left/right are the x-neighbours outside the double2, ym/yp/zm/zp the rows
// and planes beside them, bb the right-hand side.  Cells of the other
// colour or off the interior keep their bits.
__device__ __forceinline__ double2 sor_colour2(const Dims &g, int i0, int j, int k, int colour,
                                               double omega, const double2 &ctr, double left,
                                               double right, const double2 &ym,
                                               const double2 &yp, const double2 &zm,
                                               const double2 &zp, const double2 &bb, bool &any) {
    double o[2] = {ctr.x, ctr.y};
    const double vym[2] = {ym.x, ym.y}, vyp[2] = {yp.x, yp.y};
    const double vzm[2] = {zm.x, zm.y}, vzp[2] = {zp.x, zp.y}, vb[2] = {bb.x, bb.y};
    const double xm[2] = {left, ctr.x}, xp[2] = {ctr.y, right};
    any = false;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int i = i0 + e;
        if (!is_interior(g, i, j, k) || ((i + j + k) & 1) != colour) continue;
        any = true;
        double diag = 0.0, nb = 0.0;
        if (i > 1) { diag += 1.0; nb += xm[e]; }
        if (i + 2 < g.nx) { diag += 1.0; nb += xp[e]; }
        if (j > 1) { diag += 1.0; nb += vym[e]; }
        if (j + 2 < g.ny) { diag += 1.0; nb += vyp[e]; }
        if (k > 1) { diag += 1.0; nb += vzm[e]; }
        if (k + 2 < g.nz) { diag += 1.0; nb += vzp[e]; }
        const double gs = sor_quot(vb[e] + nb, diag);
        o[e] = (1.0 - omega) * o[e] + omega * gs;
    }
    return make_double2(o[0], o[1]);
}

// Red-black SOR pass of one colour on the two-wide FP64 layout, as k_rbgs4
// on the four-wide FP32 one: a cell's neighbours all have the other colour,
// which this pass leaves unchanged, so every thread reads consistent values
// and writes back its whole double2 (other-colour cells with their bits).
// Bit-identical to k_rbgs.
template <int ZU, int ST>
__global__ __launch_bounds__(kThreads) void k_rbgs2(Dims g, VTiling vt, int color, double omega,
                                                    const double *__restrict__ b,
                                                    double *__restrict__ p) {
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol2_at<ZU>(g, vt, bx, by, bz);
    const Out2<ST> o_p(p, (long)g.sxy * g.nz * 8);
    double2 pz[ZU + 2], py[ZU][2], bc[ZU];
    double xl[ZU], xr[ZU];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy;
        pz[a] = ld2<ST>(p, (base + c.off) >> 1);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        py[a][0] = ld2<ST>(p, (base + c.offym) >> 1);
        py[a][1] = ld2<ST>(p, (base + c.offyp) >> 1);
        bc[a] = ld2<ST>(b, (base + c.off) >> 1);
        xl[a] = xr[a] = 0.0;
        if (c.needL) xl[a] = ld1d<ST>(p, base + c.off - 1);
        if (c.needR) xr[a] = ld1d<ST>(p, base + c.off + 2);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const double sl = lane_prev_d(pz[a + 1].y), sr = lane_next_d(pz[a + 1].x);
        const double left = c.needL ? xl[a] : sl, right = c.needR ? xr[a] : sr;
        if (c.valid && k < g.nz) {
            bool any;
            const double2 o = sor_colour2(g, c.i0, c.j, k, color, omega, pz[a + 1], left, right,
                                          py[a][0], py[a][1], pz[a], pz[a + 2], bc[a], any);
            if (any) o_p.put((k * g.sxy + c.off) >> 1, o);
        }
    }
}

// One whole red-black SOR sweep in one launch on the two-wide FP64 layout,
// from p_in to p_out, as k_rbsor4 on the four-wide FP32 one: each thread
// forms the colour-0 values of its column and one plane past each end,
// shares its own planes through LDS (the first and last tile rows also form
// the row beyond the tile), then updates its colour-1 cells and writes its
// whole column.  The colour-0 values are what the colour-0 pass would store
// (same inputs, same arithmetic), so the sweep is bit-identical to two
// k_rbgs2 passes (and to k_rbgs).  Needs every grid row in one wave with no
// lane past its end (nx = 2 w2, w2 dividing 64).
template <int ZU, int ST>
__global__ __launch_bounds__(kThreads) void k_rbsor2(Dims g, VTiling vt, double omega,
                                                     const double *__restrict__ b,
                                                     const double *__restrict__ p_in,
                                                     double *__restrict__ p_out) {
    __shared__ double2 lds_rc[ZU][kThreads];
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol2_at<ZU>(g, vt, bx, by, bz);
    const Out2<ST> o_p(p_out, (long)g.sxy * g.nz * 8);
    const int t = threadIdx.x, ty = t / vt.txv;
    const bool first_row = ty == 0, last_row = ty == vt.ty - 1;
    const int jc = min(c.j, g.ny - 1), ic0 = c.off - jc * g.nx;
    auto row = [&](int dj) { return min(max(c.j + dj, 0), g.ny - 1) * g.nx + ic0; };
    auto plane = [&](int k) { return min(max(k, 0), g.nz - 1) * g.sxy; };
    double2 C[ZU + 4];      // row j, planes k0-2 .. k0+ZU+1
    double2 Y1[2][ZU + 2];  // rows j-1, j+1, planes k0-1 .. k0+ZU
    double2 BC[ZU + 2];     // b on row j, planes k0-1 .. k0+ZU
#pragma unroll
    for (int a = 0; a < ZU + 4; ++a) C[a] = ld2<ST>(p_in, (plane(c.k0 - 2 + a) + c.off) >> 1);
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int base = plane(c.k0 - 1 + a);
        Y1[0][a] = ld2<ST>(p_in, (base + c.offym) >> 1);
        Y1[1][a] = ld2<ST>(p_in, (base + c.offyp) >> 1);
        BC[a] = ld2<ST>(b, (base + c.off) >> 1);
    }
    // the first and last tile rows also need p two rows away and b on the
    // row beyond the tile, loaded with the rest
    const int dj_edge = first_row ? -1 : last_row ? 1 : 0;
    double2 FAR[ZU], BB[ZU];
    if (dj_edge != 0) {
        const int rfar = row(2 * dj_edge), rnear = dj_edge < 0 ? c.offym : c.offyp;
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            const int base = plane(c.k0 + a);
            FAR[a] = ld2<ST>(p_in, (base + rfar) >> 1);
            BB[a] = ld2<ST>(b, (base + rnear) >> 1);
        }
    }
    bool any;
    // colour 0 on row j, planes k0-1 .. k0+ZU; the thread's own planes to LDS
    double2 RC[ZU + 2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const double sl = lane_prev_d(C[a + 1].y), sr = lane_next_d(C[a + 1].x);
        RC[a] = sor_colour2(g, c.i0, c.j, c.k0 - 1 + a, 0, omega, C[a + 1], sl, sr, Y1[0][a],
                            Y1[1][a], C[a], C[a + 2], BC[a], any);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) lds_rc[a][t] = RC[a + 1];
    // x-neighbours on rows j-1 and j+1 (every lane takes part in the exchange)
    double yl[2][ZU], yr[2][ZU];
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            yl[s2][a] = lane_prev_d(Y1[s2][a + 1].y);
            yr[s2][a] = lane_next_d(Y1[s2][a + 1].x);
        }
    // colour 0 on the row beyond the tile (first and last tile rows only)
    double2 RE[ZU];
#pragma unroll
    for (int a = 0; a < ZU; ++a) RE[a] = make_double2(0.0, 0.0);
    auto edge_row = [&](const double2(&y)[ZU + 2], const double(&l)[ZU], const double(&r)[ZU],
                        int dj) {
#pragma unroll
        for (int a = 0; a < ZU; ++a)
            RE[a] = sor_colour2(g, c.i0, c.j + dj, c.k0 + a, 0, omega, y[a + 1], l[a], r[a],
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
        const double sl = lane_prev_d(RC[a + 1].y), sr = lane_next_d(RC[a + 1].x);
        // (indices kept in range on the rows that take RE instead)
        const double2 ym = first_row ? RE[a] : lds_rc[a][max(t - vt.txv, 0)];
        const double2 yp = last_row ? RE[a] : lds_rc[a][min(t + vt.txv, kThreads - 1)];
        if (c.valid && k < g.nz) {
            const double2 o = sor_colour2(g, c.i0, c.j, k, 1, omega, RC[a + 1], sl, sr, ym, yp,
                                          RC[a], RC[a + 2], BC[a + 1], any);
            o_p.put((k * g.sxy + c.off) >> 1, o);  // every cell: p_out holds the whole sweep
        }
    }
}

// Closes a two-wide solve whose iteration cap ended it after an even launch
// (that launch left alpha_it d_it pending): p += alpha d.  Skipped when a
// launch stopped the solve (that launch applied the pending term).
__global__ __launch_bounds__(kThreads) void k_cg_p_tail2(const CgState *st, int last_it,
                                                         const double2 *__restrict__ d,
                                                         double2 *__restrict__ p, long n2) {
    if (st->done) return;
    const double ta = st->cg_alpha[last_it & 1];
    for (long a = (long)blockIdx.x * kThreads + threadIdx.x; a < n2; a += (long)gridDim.x * kThreads)
        p[a] = axpy2(ta, d[a], p[a]);
}

}  // namespace
}  // namespace aether
