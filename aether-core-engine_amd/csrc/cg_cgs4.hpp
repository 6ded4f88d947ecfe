// cg_cgs4.hpp -- one launch per CG iteration in the single-reduction form
// (Chronopoulos-Gear) with the product w = A r carried in memory, four
// x-cells per thread.
//
// Launch `it` holds gamma_it = r.r and delta_it = r.w as partials of the
// previous launch; with beta = gamma_it / gamma_(it-1) and
// mu = delta_it - beta gamma_it / alpha_(it-1) (= d.A d), alpha = gamma_it / mu:
//   s = A d = w + beta A d_(it-1)          d <- r + beta d     p <- p + alpha d
//   r' = r - alpha s       w' = A r'       partials of r'.r', r'.w'
// The reference's loop (src/fluid.rs:463-487) produces the same iterates in
// exact arithmetic.  A d_(it-1) is not stored: the previous launch formed
// r = r_(it-1) - alpha_(it-1) A d_(it-1), so
//   s = w + (beta / alpha_(it-1)) (r_(it-1) - r)
// (s = w on the first launch).  In exact CG r_(it-1) and r are orthogonal,
// so |r_(it-1) - r| >= |r_(it-1)| and the difference loses no significant
// digits.  r' is needed only on a radius-1 neighbourhood, which each thread
// forms from the r, r_(it-1) and w it loads there (y and z halos, x edges
// at segment ends) and from its neighbour lanes (x): no LDS exchange, one
// stencil.  r and w are read on halos by other workgroups: w ping-pongs
// between two buffers and r rotates through a ring of K arrays (launch `it`
// reads r_it from ring[it % K], r_(it-1) from ring[(it - 1) % K] and writes
// ring[(it + 1) % K], which held r_(it+1-K)); d and p are updated in place.
//
// d and p are brought up to date once every K launches, by the launches
// with it mod K = K - 1.  The other launches neither load nor store them;
// the K-th rebuilds the K - 1 directions they skipped from the residuals
// those launches read, which the ring still holds at the thread's own cells
// (r_(it+1-K) in this launch's output array, read before it is
// overwritten):
//   d_j = r_j + beta_j d_(j-1)      j = it+1-K .. it   (d_0 = r_0)
//   p <- (...((p + alpha_(it+1-K) d_(it+1-K)) + ...) + alpha_it d_it
// the same FP32 operations in the same order as updating both every launch.
// A launch that stops the solve applies the p terms left pending before it;
// when the iteration cap ends it after another launch, k_cg_cgs4_tail does.
// Array passes per iteration: r, r_(it-1), w read and r', w' written (5),
// plus r_(it+1-K) .. r_(it-2), d, p read and d, p written by the K-th
// launch: 5 + (K + 2) / K on average.
//
// The kernel is bound by its instruction issue as much as by its bytes (two
// waves per SIMD, one generation, every wave loading, reducing, computing
// and storing in step), so it is written for few instructions: the FP32
// recurrences (s, r', d, p) are fused multiply-adds, the FP64 dot products
// fused multiply-adds, loads take 32-bit offsets from a scalar base, and a
// regular tiling (REG: rows that fill whole wavefront segments, dimensions
// that divide evenly) drops the edge loads, the validity tests and the
// integer divisions of the tile coordinates.
#pragma once
#include <utility>

#include "cg_vec.hpp"

namespace aether {
namespace {

// The residual ring of the one-launch CG (entries past K unused), or, for
// k_cg_cgs4_tail, the pending residuals in launch order.
struct CgRing {
    float *r[kCgRingMax];
};

// Wall masks of a regular tiling (vtiling_regular) as 0/1 factors.  With the
// row's four-cell groups one per lane, an x-neighbour of an interior cell can
// only be a wall at cell 1 of the row's first lane (mx1) and cell 2 of its
// last lane (mx2); cells 0 and 3 of those lanes are walls themselves (xin0,
// xin3).  The y factors hold for the thread's whole column.
struct RegMasks {
    float mx1, mx2, mym, myp;
    bool xin0, xin3, jin;
};

__device__ __forceinline__ RegMasks reg_masks(const Dims &g, const VCol &c) {
    RegMasks m;
    m.xin0 = c.i0 > 0;
    m.xin3 = c.i0 + 4 < g.nx;
    m.mx1 = m.xin0 ? 1.0f : 0.0f;
    m.mx2 = m.xin3 ? 1.0f : 0.0f;
    m.mym = c.j > 1 ? 1.0f : 0.0f;
    m.myp = c.j + 2 < g.ny ? 1.0f : 0.0f;
    m.jin = c.j >= 1 && c.j <= g.ny - 2;
    return m;
}

// lap4 for a regular tiling: the same sum of (x_c - x_nb) over the valid
// neighbours in the same order (x-, x+, y-, y+, z-, z+), with each term that
// can be masked added as a multiply-add by its 0/1 factor.  Equal to lap4 up
// to the sign of a zero result.
__device__ __forceinline__ void lap4_reg(const Dims &g, const RegMasks &m, int k, const float4 &x,
                                         float xl, float xr, const float4 &ym, const float4 &yp,
                                         const float4 &zm, const float4 &zp, float (&q)[4]) {
    const float mzm = k > 1 ? 1.0f : 0.0f, mzp = k + 2 < g.nz ? 1.0f : 0.0f;
    const bool kin = k >= 1 && k <= g.nz - 2 && m.jin;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const float xc = f4(x, e);
        const float xm = e == 0 ? xl : f4(x, e - 1), xp = e == 3 ? xr : f4(x, e + 1);
        float acc = e == 1 ? m.mx1 * (xc - xm) : xc - xm;
        acc = e == 2 ? __builtin_fmaf(m.mx2, xc - xp, acc) : acc + (xc - xp);
        acc = __builtin_fmaf(m.mym, xc - f4(ym, e), acc);
        acc = __builtin_fmaf(m.myp, xc - f4(yp, e), acc);
        acc = __builtin_fmaf(mzm, xc - f4(zm, e), acc);
        acc = __builtin_fmaf(mzp, xc - f4(zp, e), acc);
        const bool in = kin && (e == 0 ? m.xin0 : e == 3 ? m.xin3 : true);
        q[e] = in ? acc : 0.0f;
    }
}

// Applies at cell `oc` the p terms of the M launches before launch `it`
// (it mod K = PH), in launch order: rebuilds d_(it-M) .. d_(it-1) from
// dd = d_(it-M-1) (unused when it == M: d_0 = r_0) and the ring's residuals,
// and returns d_(it-1) in dd and the updated p in pv.  fa/fb[m]: alpha and
// beta of launch it - m.
template <int K, int PH, int M, int ST>
__device__ __forceinline__ void cgs4_pending(const CgRing &ring, int it, int oc,
                                             const float (&fa)[K], const float (&fb)[K],
                                             float4 &dd, float4 &pv) {
#pragma unroll
    for (int m = M; m >= 1; --m) {
        const float4 rj = ld4<ST>(ring.r[(PH - m + K) % K], oc);
        dd = it == m ? rj : fma4(fb[m], dd, rj);
        pv = fma4(fa[m], dd, pv);
    }
}

// RG: 0 any tiling; 1 a regular tiling (vtiling_regular); 2 a regular tiling
// with two rows per wavefront (32 float4 lanes per row), where each lane takes
// the y-neighbour row inside its wavefront from the other half of the wave and
// loads only the row outside it.
// The work of one workgroup on the tile `c`; `slot` is its partial-sum slot
// (0 leads: it publishes the CG scalars).
template <int ZU, bool FIRST, int NQ, int ST, int K, int PH, int RG>
__device__ __forceinline__ void body_cg_cgs4(
    const VCol &c, int slot, const Dims &g, int it, double rel_tol2, int nb_in,
    const double *__restrict__ pg_in, const double *__restrict__ pd_in,
    const double *__restrict__ pbb, double *__restrict__ pg_out, double *__restrict__ pd_out,
    CgState *st, const CgRing &ring, const float *__restrict__ w_in, float *__restrict__ w_out,
    float *__restrict__ d, float *__restrict__ p, int *done_mirror, int solve_id) {
    constexpr bool REG = RG != 0;
    // K >= 3: the array a launch overwrites is neither its r_it nor its r_(it-1)
    static_assert(K >= 3 && K <= kCgRingMax && PH >= 0 && PH < K && (!FIRST || PH == 0),
                  "one-launch CG cycle");
    constexpr int kMask = kCgRingMax - 1;
    // separate slots for each reduction, so none needs a trailing barrier
    __shared__ double lds[8], lds_b[8], lds_out[8];
    __shared__ int flag_lds[1];
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> eg, ed, eb;
    eg.load(pg_in, nb_in);
    ed.load(pd_in, nb_in);
    if (FIRST) eb.load(pbb, nb_in);
    double gprev = 1.0, aprev = 1.0, b2 = 0.0;
    if (!FIRST) {
        gprev = st->cg_gamma[(it - 1) & 1];
        aprev = st->cg_alpha[(it - 1) & kMask];
        b2 = st->b2;
    }
    // alpha and beta of the launches whose p terms are pending: fa/fb[m] of
    // launch it - m, m = 1 .. PH
    float fa[K], fb[K];
#pragma unroll
    for (int m = 0; m < K; ++m) fa[m] = fb[m] = 0.0f;
#pragma unroll
    for (int m = 1; m <= PH; ++m) {
        fa[m] = (float)st->cg_alpha[(it - m) & kMask];
        fb[m] = (float)st->cg_beta[(it - m) & kMask];
    }
    const float *__restrict__ r_in = ring.r[PH];
    float *__restrict__ r_out = ring.r[(PH + 1) % K];
    const float *__restrict__ r_prev = ring.r[(PH + K - 1) % K];
    const long nbytes = (long)g.sxy * g.nz * 4;
    const Out4<ST> o_r(r_out, nbytes), o_w(w_out, nbytes), o_d(d, nbytes), o_p(p, nbytes);
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // r, w and r_(it-1) on the radius-1 neighbourhood (z column, y rows, x
    // edges), issued plane by plane: z plane a, then the y rows of plane
    // a - 1, so that the stencil of a plane waits only for the loads it reads
    float4 rz[ZU + 2], wz[ZU + 2], qz[ZU + 2], ry[ZU][2], wy[ZU][2], qy[ZU][2];
    float er[ZU][2], ew[ZU][2], eq[ZU][2];
    const bool hi = c.lane >= 32;  // RG 2: the wavefront's second row
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int o = (min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy + c.off) >> 2;
        rz[a] = ld4<ST>(r_in, o);
        wz[a] = ld4<ST>(w_in, o);
        qz[a] = FIRST ? z4 : ld4<ST>(r_prev, o);
        if (a < 1 || a > ZU) continue;
        const int b = a - 1;
        const int base = min(c.k0 + b, g.nz - 1) * g.sxy;
        if constexpr (RG == 2) {  // only the row outside the wavefront: [0]
            const int oo = (base + (hi ? c.offyp : c.offym)) >> 2;
            ry[b][0] = ld4<ST>(r_in, oo);
            wy[b][0] = ld4<ST>(w_in, oo);
            qy[b][0] = FIRST ? z4 : ld4<ST>(r_prev, oo);
            ry[b][1] = wy[b][1] = qy[b][1] = z4;
            continue;
        }
        const int om = (base + c.offym) >> 2, op = (base + c.offyp) >> 2;
        ry[b][0] = ld4<ST>(r_in, om);
        ry[b][1] = ld4<ST>(r_in, op);
        wy[b][0] = ld4<ST>(w_in, om);
        wy[b][1] = ld4<ST>(w_in, op);
        qy[b][0] = FIRST ? z4 : ld4<ST>(r_prev, om);
        qy[b][1] = FIRST ? z4 : ld4<ST>(r_prev, op);
#pragma unroll
        for (int e = 0; e < 2; ++e) er[b][e] = ew[b][e] = eq[b][e] = 0.0f;
        if (!REG && c.needL) {
            er[b][0] = ld1<ST>(r_in, base + c.off - 1);
            ew[b][0] = ld1<ST>(w_in, base + c.off - 1);
            if (!FIRST) eq[b][0] = ld1<ST>(r_prev, base + c.off - 1);
        }
        if (!REG && c.needR) {
            er[b][1] = ld1<ST>(r_in, base + c.off + 4);
            ew[b][1] = ld1<ST>(w_in, base + c.off + 4);
            if (!FIRST) eq[b][1] = ld1<ST>(r_prev, base + c.off + 4);
        }
    }

    // ---- scalars (every workgroup reaches the same decision) -------------
    // The solve's stop word holds (stopping launch + 1).  Only launches after
    // the stopping one take the early exit: a workgroup of the stopping
    // launch may already see that launch's word (written mid-launch by
    // workgroup 0), and it must still reach the stop below, which applies
    // the pending p terms.
    if (threadIdx.x == 0) flag_lds[0] = flag_reg != 0 && flag_reg <= it;
    double tot[2] = {eg.local(pg_in, nb_in), ed.local(pd_in, nb_in)};
    block_sum<2, false>(tot, lds);  // its barrier publishes flag_lds
    if (flag_lds[0]) return;
    const double gam = tot[0], del = tot[1];
    const bool lead = slot == 0 && threadIdx.x == 0;
    double beta = 0.0, mu = del, tcd = 0.0;  // tcd: beta / alpha_(it-1)
    bool stop;
    if (FIRST) {
        b2 = eb.total(pbb, nb_in, lds_b);
        stop = !(gam > 1e-24) || (rel_tol2 > 0.0 && gam <= rel_tol2 * b2);
        if (lead) {
            st->b2 = b2;
            st->r2_init = gam;
        }
    } else {
        stop = gam < 1e-20 || (rel_tol2 > 0.0 && gam <= rel_tol2 * b2);
        beta = gam / gprev;
        tcd = beta / aprev;
        mu = del - tcd * gam;
    }
    if (!stop && fabs(mu) < 1e-28) stop = true;
    if (stop) {
        if (lead) {
            st->done = it + 1;
            st->iters = it;
            st->r2_final = gam;
            st->converged = 1;
            *done_mirror = solve_id;
        }
        if (PH >= 1) {  // the terms of the launches since the last K-th one
#pragma unroll
            for (int a = 0; a < ZU; ++a) {
                const int k = c.k0 + a;
                if (REG || (c.valid && k < g.nz)) {
                    const int oc = (k * g.sxy + c.off) >> 2;
                    float4 dd = ld4<ST>(d, oc), pv = ld4<ST>(p, oc);
                    cgs4_pending<K, PH, PH, ST>(ring, it, oc, fa, fb, dd, pv);
                    o_p.put(oc, pv);
                }
            }
        }
        return;
    }
    const double alpha = gam / mu;
    if (lead) {
        st->cg_gamma[it & 1] = gam;
        st->cg_alpha[it & kMask] = alpha;
        st->cg_beta[it & kMask] = beta;
    }
    const float tb = (float)beta, ta = (float)alpha;
    // beta / alpha_(it-1): scales r_(it-1) - r into beta A d_(it-1)
    const float tc = (float)tcd;

    // ---- s = w + (beta / alpha_(it-1)) (r_(it-1) - r) and r' = r - alpha s
    // on the radius-1 neighbourhood
    float4 rn[ZU + 2], rny[ZU][2];
    float rne[ZU][2];
    auto r_new = [&](const float4 &rv, const float4 &wv, const float4 &qv) {
        const float4 sv = FIRST ? wv : fma4(tc, sub4(qv, rv), wv);
        return fma4(-ta, sv, rv);
    };
    auto r_new_y = [&](int a) {  // the y rows (and row-segment edges) of plane a
#pragma unroll
        for (int e = 0; e < (RG == 2 ? 1 : 2); ++e) {
            rny[a][e] = r_new(ry[a][e], wy[a][e], qy[a][e]);
            if (!REG) {
                const float se =
                    FIRST ? ew[a][e] : __builtin_fmaf(tc, eq[a][e] - er[a][e], ew[a][e]);
                rne[a][e] = __builtin_fmaf(-ta, se, er[a][e]);
            } else {
                rne[a][e] = 0.0f;
            }
        }
    };

    // ---- w' = A r' and the partials of r'.r', r'.w' on plane a ------------
    RegMasks rm = reg_masks(g, c);
    if constexpr (RG == 2) {  // the stencil's y terms in the order (outside row, inside row)
        const float fm = rm.mym, fp = rm.myp;
        rm.mym = hi ? fp : fm;
        rm.myp = hi ? fm : fp;
    }
    double acc[2] = {0.0, 0.0};
    auto plane = [&](int a) {
        const int k = c.k0 + a;
        // x-neighbours from the adjacent lanes by whole-wave shifts (lanes 0
        // and 63 sit at a row start or end, or take the edge loads)
        const float sl = lane_prev(rn[a + 1].w), sr = lane_next(rn[a + 1].x);
        const float xl = !REG && c.needL ? rne[a][0] : sl;
        const float xrv = !REG && c.needR ? rne[a][1] : sr;
        if (REG || (c.valid && k < g.nz)) {
            float q[4];
            if constexpr (RG == 2) {
                const float4 &x = rn[a + 1];
                const float4 yin = make_float4(other_half(x.x, hi), other_half(x.y, hi),
                                               other_half(x.z, hi), other_half(x.w, hi));
                lap4_reg(g, rm, k, x, xl, xrv, rny[a][0], yin, rn[a], rn[a + 2], q);
            } else if constexpr (REG)
                lap4_reg(g, rm, k, rn[a + 1], xl, xrv, rny[a][0], rny[a][1], rn[a], rn[a + 2], q);
            else
                lap4(g, c, k, rn[a + 1], xl, xrv, rny[a][0], rny[a][1], rn[a], rn[a + 2], q);
            const float4 wn = make_float4(q[0], q[1], q[2], q[3]);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double rv = (double)f4(rn[a + 1], e);
                acc[0] = __builtin_fma(rv, rv, acc[0]);
                acc[1] = __builtin_fma(rv, (double)q[e], acc[1]);
            }
            const int id4 = (k * g.sxy + c.off) >> 2;
            o_r.put(id4, rn[a + 1]);
            o_w.put(id4, wn);
        }
    };

    // ---- every K-th launch: d_(it+1-K) .. d_it and p, finished and stored
    // at once (the ring's residuals are those of the K - 1 launches before;
    // r_(it+1-K) sits in this launch's output array and is read here, before
    // the planes below overwrite it).  r' is formed on the whole
    // neighbourhood first, so that the halo registers are dead while the
    // ring's residuals, d and p are in flight; the planes' r' and w' are
    // stored last, so that the next launch finds them in L2.
    auto rebuild_dp = [&]() {
        float4 dd[ZU], pv[ZU];
        int oc[ZU];
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            const int k = REG ? c.k0 + a : min(c.k0 + a, g.nz - 1);
            oc[a] = (k * g.sxy + c.off) >> 2;
            dd[a] = ld4<ST>(d, oc[a]);
            pv[a] = ld4<ST>(p, oc[a]);
        }
        // launches it+1-K .. it-1 in order: d_j = r_j + beta_j d_(j-1)
        // (d_0 = r_0), p += alpha_j d_j; launch it - m read ring[(PH - m) % K].
        // Only the solve's first K-th launch starts from d_0 = r_0; that one
        // decision is taken before the loop, so that the loop's loads share
        // one basic block and are all in flight together.
        float4 rold[ZU];
#pragma unroll
        for (int a = 0; a < ZU; ++a) rold[a] = ld4<ST>(r_out, oc[a]);
        if (it == K - 1) {
#pragma unroll
            for (int a = 0; a < ZU; ++a) dd[a] = rold[a];
        } else {
#pragma unroll
            for (int a = 0; a < ZU; ++a) dd[a] = fma4(fb[K - 1], dd[a], rold[a]);
        }
#pragma unroll
        for (int a = 0; a < ZU; ++a) pv[a] = fma4(fa[K - 1], dd[a], pv[a]);
#pragma unroll
        for (int m = K - 2; m >= 1; --m) {
#pragma unroll
            for (int a = 0; a < ZU; ++a) {
                const float4 rj = ld4<ST>(ring.r[(PH - m + K) % K], oc[a]);
                dd[a] = fma4(fb[m], dd[a], rj);
                pv[a] = fma4(fa[m], dd[a], pv[a]);
            }
        }
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            dd[a] = fma4(tb, dd[a], rz[a + 1]);  // d_it
            pv[a] = fma4(ta, dd[a], pv[a]);
            if (REG || (c.valid && c.k0 + a < g.nz)) {
                o_d.put(oc[a], dd[a]);
                o_p.put(oc[a], pv[a]);
            }
        }
    };

    if constexpr (PH != K - 1) {
        // plane by plane, in the order the loads were issued
        rn[0] = r_new(rz[0], wz[0], qz[0]);
        rn[1] = r_new(rz[1], wz[1], qz[1]);
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            rn[a + 2] = r_new(rz[a + 2], wz[a + 2], qz[a + 2]);
            r_new_y(a);
            plane(a);
            __builtin_amdgcn_sched_barrier(0);  // this plane's stores go out now
        }
    } else {
#pragma unroll
        for (int a = 0; a < ZU + 2; ++a) rn[a] = r_new(rz[a], wz[a], qz[a]);
#pragma unroll
        for (int a = 0; a < ZU; ++a) r_new_y(a);
        __builtin_amdgcn_sched_barrier(0);
        rebuild_dp();
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int a = 0; a < ZU; ++a) plane(a);
    }
    block_sum<2, false>(acc, lds_out);
    if (threadIdx.x == 0) {
        pg_out[slot] = acc[0];
        pd_out[slot] = acc[1];
    }
}

template <int ZU, bool FIRST, int NQ, int ST, int K, int PH, int RG>
__global__ __launch_bounds__(kThreads) void k_cg_cgs4(
    Dims g, VTiling vt, int it, double rel_tol2, int nb_in, const double *__restrict__ pg_in,
    const double *__restrict__ pd_in, const double *__restrict__ pbb, double *__restrict__ pg_out,
    double *__restrict__ pd_out, CgState *st, CgRing ring, const float *__restrict__ w_in,
    float *__restrict__ w_out, float *__restrict__ d, float *__restrict__ p, int *done_mirror,
    int solve_id) {
    VCol c;
    if constexpr (RG != 0) {
        c = vcol_regular<ZU>(g, vt);
    } else {
        int bx, by, bz;
        vtile_coords(vt, bx, by, bz);
        c = vcol_at<ZU>(g, vt, bx, by, bz);
    }
    body_cg_cgs4<ZU, FIRST, NQ, ST, K, PH, RG>(c, blockIdx.x, g, it, rel_tol2, nb_in, pg_in,
                                              pd_in, pbb, pg_out, pd_out, st, ring, w_in, w_out,
                                              d, p, done_mirror, solve_id);
}

// Closes a solve whose iteration cap ended it after launch `last_it` of
// k_cg_cgs4 when that was not a K-th launch (last_it mod K < K - 1): it
// rebuilds the directions since the last K-th launch and applies their
// pending p terms, in the order the K-th launch would have.  pend.r[q]: the
// residual that launch last_it - last_it mod K + q read, q = 0 .. last_it
// mod K.  Skipped when a launch stopped the solve (that launch applied what
// was pending).
template <int K>
__global__ __launch_bounds__(kThreads) void k_cg_cgs4_tail(const CgState *st, int last_it,
                                                           CgRing pend,
                                                           const float4 *__restrict__ d,
                                                           float4 *__restrict__ p, long n4) {
    constexpr int kMask = kCgRingMax - 1;
    if (st->done) return;
    const int L = last_it, npend = L % K + 1, j0 = L - L % K;
    float fa[K], fb[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        fa[q] = q < npend ? (float)st->cg_alpha[(j0 + q) & kMask] : 0.0f;
        fb[q] = q < npend ? (float)st->cg_beta[(j0 + q) & kMask] : 0.0f;
    }
    for (long a = (long)blockIdx.x * kThreads + threadIdx.x; a < n4;
         a += (long)gridDim.x * kThreads) {
        float4 dd = d[a], pv = p[a];
#pragma unroll
        for (int q = 0; q < K - 1; ++q) {
            if (q < npend) {
                const float4 rj = reinterpret_cast<const float4 *>(pend.r[q])[a];
                dd = j0 + q == 0 ? rj : fma4(fb[q], dd, rj);
                pv = fma4(fa[q], dd, pv);
            }
        }
        p[a] = pv;
    }
}

// ---- host-side selection of the instantiations --------------------------

// Launches per d/p update of the one-launch four-wide CG: 6.25 passes per
// iteration, and a 200-iteration cap needs no tail launch (a cycle of 3: 6.67
// passes and a tail; the 128^3 step measured ~1% slower with 3 or 4).
constexpr int kCgsCycle = 8;

template <int NQ, int ST, int RG, int... PH>
inline const void *cgs4_phase(int ph, std::integer_sequence<int, PH...>) {
    const void *fns[] = {(const void *)k_cg_cgs4<4, false, NQ, ST, kCgsCycle, PH, RG>...};
    return fns[ph];
}

template <int NQ, int ST, int RG>
inline const void *cgs4_pick(bool first, int ph) {
    if (first) return (const void *)k_cg_cgs4<4, true, NQ, ST, kCgsCycle, 0, RG>;
    return cgs4_phase<NQ, ST, RG>(ph, std::make_integer_sequence<int, kCgsCycle>{});
}

// The k_cg_cgs4 instantiation for launch phase ph (= it mod kCgsCycle).
// nparts: partial sums read or written per launch (2, 4 or 16 per thread); wt:
// write-through stores and 32-bit load offsets (arrays under 2 GiB), else
// plain stores; rg: the tiling class (RG of k_cg_cgs4; 1 and 2 need wt and at
// most 1024 partials).
inline const void *cgs4_kernel(bool first, int ph, int nparts, bool wt, int rg) {
    const bool wide = nparts > 4 * kThreads;
    if (wt && rg != 0 && !wide) {
        const bool two = nparts <= 2 * kThreads;
        if (rg == 2) return two ? cgs4_pick<2, kStWT, 2>(first, ph) : cgs4_pick<4, kStWT, 2>(first, ph);
        return two ? cgs4_pick<2, kStWT, 1>(first, ph) : cgs4_pick<4, kStWT, 1>(first, ph);
    }
    if (wide)
        return wt ? cgs4_pick<16, kStWT, 0>(first, ph) : cgs4_pick<16, kStPlain, 0>(first, ph);
    return wt ? cgs4_pick<4, kStWT, 0>(first, ph) : cgs4_pick<4, kStPlain, 0>(first, ph);
}

inline void cgs4_tail_launch(unsigned nb, hipStream_t st, const CgState *cg, int last,
                             const CgRing &pend, const float4 *d, float4 *p, long n4) {
    k_cg_cgs4_tail<kCgsCycle><<<nb, kThreads, 0, st>>>(cg, last, pend, d, p, n4);
}

}  // namespace
}  // namespace aether
