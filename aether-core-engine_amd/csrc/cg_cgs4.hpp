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
// between two buffers and r rotates through three (launch `it` reads
// R[it % 3], finds r_(it-1) in R[(it + 2) % 3] and writes R[(it + 1) % 3],
// which held r_(it-2)); d and p are updated in place.
//
// d and p are brought up to date every third launch (PH = it mod 3 = 2).
// The other launches neither load nor store them; the third rebuilds the
// two directions they skipped from the residuals those launches read, which
// the rotation still holds at the thread's own cells (r_(it-2) in this
// launch's output array, read before it is overwritten; r_(it-1)):
//   d_(it-2) = r_(it-2) + beta_(it-2) d_(it-3)      (d_0 = r_0)
//   d_(it-1) = r_(it-1) + beta_(it-1) d_(it-2)
//   d_it     = r_it + beta_it d_(it-1)
//   p <- ((p + alpha_(it-2) d_(it-2)) + alpha_(it-1) d_(it-1)) + alpha_it d_it
// the same FP32 operations in the same order as updating both every launch.
// A launch that stops the solve applies the p terms left pending before it;
// when the iteration cap ends it after another launch, k_cg_cgs4_tail does.
// Array passes per iteration: r, r_(it-1), w read and r', w' written (5),
// plus r_(it-2), d, p read and d, p written every third launch: 6.67 on
// average (10 on every third launch, 5 on the others).
#pragma once
#include "cg_vec.hpp"

namespace aether {

template <int ZU, bool FIRST, int NQ, int ST, int PH>
__global__ __launch_bounds__(kThreads) void k_cg_cgs4(
    Dims g, VTiling vt, int it, double rel_tol2, int nb_in, const double *__restrict__ pg_in,
    const double *__restrict__ pd_in, const double *__restrict__ pbb, double *__restrict__ pg_out,
    double *__restrict__ pd_out, CgState *st, const float *__restrict__ r_in,
    float *__restrict__ r_out, const float *__restrict__ r_prev,
    const float *__restrict__ w_in, float *__restrict__ w_out, float *__restrict__ d,
    float *__restrict__ p, int *done_mirror, int solve_id) {
    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const VCol c = vcol_at<ZU>(g, vt, bx, by, bz);
    // separate slots for each reduction, so none needs a trailing barrier
    __shared__ double lds[8], lds_b[8], lds_out[8];
    __shared__ int flag_lds[1];
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> eg, ed, eb;
    eg.load(pg_in, nb_in);
    ed.load(pd_in, nb_in);
    if (FIRST) eb.load(pbb, nb_in);
    double gprev = 1.0, aprev = 1.0, aprev2 = 0.0, bprev = 0.0, bprev2 = 0.0, b2 = 0.0;
    if (!FIRST) {
        gprev = st->cg_gamma[(it - 1) & 1];
        aprev = st->cg_alpha[(it - 1) & 3];
        b2 = st->b2;
    }
    if (PH >= 1) bprev = st->cg_beta[(it - 1) & 3];
    if (PH == 2) {
        aprev2 = st->cg_alpha[(it - 2) & 3];
        bprev2 = st->cg_beta[(it - 2) & 3];
    }
    const float4 *r4 = reinterpret_cast<const float4 *>(r_in);
    // r_(it-2) (read at this thread's own cells before they are overwritten)
    // and r_(it-1)
    const float4 *ro2 = reinterpret_cast<const float4 *>(r_out);
    const float4 *ro1 = reinterpret_cast<const float4 *>(r_prev);
    const float4 *w4 = reinterpret_cast<const float4 *>(w_in);
    float4 *d4 = reinterpret_cast<float4 *>(d);
    float4 *p4 = reinterpret_cast<float4 *>(p);
    const long nbytes = (long)g.sxy * g.nz * 4;
    const Out4<ST> o_r(r_out, nbytes), o_w(w_out, nbytes), o_d(d, nbytes), o_p(p, nbytes);
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // r, w and r_(it-1) on the radius-1 neighbourhood (z column, y rows, x edges)
    float4 rz[ZU + 2], wz[ZU + 2], qz[ZU + 2], ry[ZU][2], wy[ZU][2], qy[ZU][2];
    float er[ZU][2], ew[ZU][2], eq[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int o = (min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy + c.off) >> 2;
        rz[a] = r4[o];
        wz[a] = w4[o];
        qz[a] = FIRST ? z4 : ro1[o];
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        const int om = (base + c.offym) >> 2, op = (base + c.offyp) >> 2;
        ry[a][0] = r4[om];
        ry[a][1] = r4[op];
        wy[a][0] = w4[om];
        wy[a][1] = w4[op];
        qy[a][0] = FIRST ? z4 : ro1[om];
        qy[a][1] = FIRST ? z4 : ro1[op];
#pragma unroll
        for (int e = 0; e < 2; ++e) er[a][e] = ew[a][e] = eq[a][e] = 0.0f;
        if (c.needL) {
            er[a][0] = r_in[base + c.off - 1];
            ew[a][0] = w_in[base + c.off - 1];
            if (!FIRST) eq[a][0] = r_prev[base + c.off - 1];
        }
        if (c.needR) {
            er[a][1] = r_in[base + c.off + 4];
            ew[a][1] = w_in[base + c.off + 4];
            if (!FIRST) eq[a][1] = r_prev[base + c.off + 4];
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
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    double beta = 0.0, mu = del;
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
        mu = del - beta * gam / aprev;
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
        if (PH >= 1) {  // the terms of the launches since the last third one
            const float tap = (float)aprev, tbp = (float)bprev, tap2 = (float)aprev2,
                        tbp2 = (float)bprev2;
#pragma unroll
            for (int a = 0; a < ZU; ++a) {
                const int k = c.k0 + a;
                if (c.valid && k < g.nz) {
                    const int oc = (k * g.sxy + c.off) >> 2;
                    const float4 dv = d4[oc], r1 = ro1[oc];
                    if (PH == 1) {  // alpha_(it-1) d_(it-1), d_(it-2) stored
                        const float4 d1 = it == 1 ? r1 : axpy4(r1, tbp, dv);
                        o_p.put(oc, axpy4(p4[oc], tap, d1));
                    } else {  // and alpha_(it-2) d_(it-2), d_(it-3) stored
                        const float4 r2 = ro2[oc];
                        const float4 d2 = it == 2 ? r2 : axpy4(r2, tbp2, dv);
                        const float4 d1 = axpy4(r1, tbp, d2);
                        o_p.put(oc, axpy4(axpy4(p4[oc], tap2, d2), tap, d1));
                    }
                }
            }
        }
        return;
    }
    const double alpha = gam / mu;
    if (lead) {
        st->cg_gamma[it & 1] = gam;
        st->cg_alpha[it & 3] = alpha;
        st->cg_beta[it & 3] = beta;
    }
    const float tb = (float)beta, ta = (float)alpha, tap = (float)aprev, tbp = (float)bprev,
                tap2 = (float)aprev2, tbp2 = (float)bprev2;
    // beta / alpha_(it-1): scales r_(it-1) - r into beta A d_(it-1)
    const float tc = FIRST ? 0.0f : (float)(beta / aprev);

    // ---- s = w + (beta / alpha_(it-1)) (r_(it-1) - r) and r' = r - alpha s
    // on the radius-1 neighbourhood
    float4 rn[ZU + 2], rny[ZU][2];
    float rne[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const float4 sv = FIRST ? wz[a] : axpy4(wz[a], tc, sub4(qz[a], rz[a]));
        rn[a] = make_float4(rz[a].x - ta * sv.x, rz[a].y - ta * sv.y, rz[a].z - ta * sv.z,
                            rz[a].w - ta * sv.w);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float4 sv = FIRST ? wy[a][e] : axpy4(wy[a][e], tc, sub4(qy[a][e], ry[a][e]));
            rny[a][e] = make_float4(ry[a][e].x - ta * sv.x, ry[a][e].y - ta * sv.y,
                                    ry[a][e].z - ta * sv.z, ry[a][e].w - ta * sv.w);
            const float se = FIRST ? ew[a][e] : ew[a][e] + tc * (eq[a][e] - er[a][e]);
            rne[a][e] = er[a][e] - ta * se;
        }
    }

    // ---- every third launch: d_(it-2), d_(it-1), d_it and p, finished and
    // stored at once.  Loaded only now, after the r, w, s registers above
    // are consumed, and held by nothing below, so that the launch keeps two
    // waves per SIMD.
    if (PH == 2) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            const int k = c.k0 + a;
            const int oc = (min(k, g.nz - 1) * g.sxy + c.off) >> 2;
            const float4 r2 = ro2[oc], r1 = ro1[oc], dv = d4[oc], pv = p4[oc];
            const float4 d2 = it == 2 ? r2 : axpy4(r2, tbp2, dv);  // d_(it-2); d_0 = r_0
            const float4 d1 = axpy4(r1, tbp, d2);                 // d_(it-1)
            const float4 dn = axpy4(rz[a + 1], tb, d1);            // d_it
            if (c.valid && k < g.nz) {
                o_d.put(oc, dn);
                o_p.put(oc, axpy4(axpy4(axpy4(pv, tap2, d2), tap, d1), ta, dn));
            }
        }
    }

    // ---- w' = A r', d, p and the partials of r'.r', r'.w' ----------------
    const int srcL = (c.lane + 63) & 63, srcR = (c.lane + 1) & 63;
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const float sl = lane_value(rn[a + 1].w, srcL), sr = lane_value(rn[a + 1].x, srcR);
        const float xl = c.needL ? rne[a][0] : sl;
        const float xrv = c.needR ? rne[a][1] : sr;
        if (c.valid && k < g.nz) {
            float q[4];
            lap4(g, c, k, rn[a + 1], xl, xrv, rny[a][0], rny[a][1], rn[a], rn[a + 2], q);
            const float4 wn = make_float4(q[0], q[1], q[2], q[3]);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double rv = (double)f4(rn[a + 1], e);
                acc[0] += rv * rv;
                acc[1] += rv * (double)q[e];
            }
            const int id4 = (k * g.sxy + c.off) >> 2;
            o_r.put(id4, rn[a + 1]);
            o_w.put(id4, wn);
        }
    }
    block_sum<2, false>(acc, lds_out);
    if (threadIdx.x == 0) {
        pg_out[blockIdx.x] = acc[0];
        pd_out[blockIdx.x] = acc[1];
    }
}

// Closes a solve whose iteration cap ended it after launch `last_it` of
// k_cg_cgs4 when that was not a third launch (last_it mod 3 = 0 or 1): it
// rebuilds the directions since the last third launch and applies their
// pending p terms, in the order the third launch would have.  r_lm1 and
// r_l: the residuals launches last_it - 1 and last_it read.  Skipped when a
// launch stopped the solve (that launch applied what was pending).
__global__ __launch_bounds__(kThreads) void k_cg_cgs4_tail(const CgState *st, int last_it,
                                                           const float4 *__restrict__ r_lm1,
                                                           const float4 *__restrict__ r_l,
                                                           const float4 *__restrict__ d,
                                                           float4 *__restrict__ p, long n4) {
    if (st->done) return;
    const int L = last_it;
    const float ta = (float)st->cg_alpha[L & 3], tb = (float)st->cg_beta[L & 3];
    const float ta1 = L >= 1 ? (float)st->cg_alpha[(L - 1) & 3] : 0.0f;
    const float tb1 = L >= 1 ? (float)st->cg_beta[(L - 1) & 3] : 0.0f;
    for (long a = (long)blockIdx.x * kThreads + threadIdx.x; a < n4;
         a += (long)gridDim.x * kThreads) {
        if (L % 3 == 0) {  // alpha_L d_L; d_(L-1) stored
            const float4 dl = L == 0 ? r_l[a] : axpy4(r_l[a], tb, d[a]);
            p[a] = axpy4(p[a], ta, dl);
        } else {  // alpha_(L-1) d_(L-1) + alpha_L d_L; d_(L-2) stored
            const float4 d1 = L == 1 ? r_lm1[a] : axpy4(r_lm1[a], tb1, d[a]);
            const float4 dl = axpy4(r_l[a], tb, d1);
            p[a] = axpy4(axpy4(p[a], ta1, d1), ta, dl);
        }
    }
}

}  // namespace aether
