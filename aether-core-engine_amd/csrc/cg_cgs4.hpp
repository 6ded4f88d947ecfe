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

template <int ZU, bool FIRST, int NQ, int ST, int K, int PH, bool REG>
__global__ __launch_bounds__(kThreads) void k_cg_cgs4(
    Dims g, VTiling vt, int it, double rel_tol2, int nb_in, const double *__restrict__ pg_in,
    const double *__restrict__ pd_in, const double *__restrict__ pbb, double *__restrict__ pg_out,
    double *__restrict__ pd_out, CgState *st, CgRing ring, const float *__restrict__ w_in,
    float *__restrict__ w_out, float *__restrict__ d, float *__restrict__ p, int *done_mirror,
    int solve_id) {
    // K >= 3: the array a launch overwrites is neither its r_it nor its r_(it-1)
    static_assert(K >= 3 && K <= kCgRingMax && PH >= 0 && PH < K && (!FIRST || PH == 0),
                  "one-launch CG cycle");
    constexpr int kMask = kCgRingMax - 1;
    VCol c;
    if constexpr (REG) {
        c = vcol_regular<ZU>(g, vt);
    } else {
        int bx, by, bz;
        vtile_coords(vt, bx, by, bz);
        c = vcol_at<ZU>(g, vt, bx, by, bz);
    }
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
    // r, w and r_(it-1) on the radius-1 neighbourhood (z column, y rows, x edges)
    float4 rz[ZU + 2], wz[ZU + 2], qz[ZU + 2], ry[ZU][2], wy[ZU][2], qy[ZU][2];
    float er[ZU][2], ew[ZU][2], eq[ZU][2];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const int o = (min(max(c.k0 - 1 + a, 0), g.nz - 1) * g.sxy + c.off) >> 2;
        rz[a] = ld4<ST>(r_in, o);
        wz[a] = ld4<ST>(w_in, o);
        qz[a] = FIRST ? z4 : ld4<ST>(r_prev, o);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int base = min(c.k0 + a, g.nz - 1) * g.sxy;
        const int om = (base + c.offym) >> 2, op = (base + c.offyp) >> 2;
        ry[a][0] = ld4<ST>(r_in, om);
        ry[a][1] = ld4<ST>(r_in, op);
        wy[a][0] = ld4<ST>(w_in, om);
        wy[a][1] = ld4<ST>(w_in, op);
        qy[a][0] = FIRST ? z4 : ld4<ST>(r_prev, om);
        qy[a][1] = FIRST ? z4 : ld4<ST>(r_prev, op);
#pragma unroll
        for (int e = 0; e < 2; ++e) er[a][e] = ew[a][e] = eq[a][e] = 0.0f;
        if (!REG && c.needL) {
            er[a][0] = ld1<ST>(r_in, base + c.off - 1);
            ew[a][0] = ld1<ST>(w_in, base + c.off - 1);
            if (!FIRST) eq[a][0] = ld1<ST>(r_prev, base + c.off - 1);
        }
        if (!REG && c.needR) {
            er[a][1] = ld1<ST>(r_in, base + c.off + 4);
            ew[a][1] = ld1<ST>(w_in, base + c.off + 4);
            if (!FIRST) eq[a][1] = ld1<ST>(r_prev, base + c.off + 4);
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
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) {
        const float4 sv = FIRST ? wz[a] : fma4(tc, sub4(qz[a], rz[a]), wz[a]);
        rn[a] = fma4(-ta, sv, rz[a]);
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const float4 sv = FIRST ? wy[a][e] : fma4(tc, sub4(qy[a][e], ry[a][e]), wy[a][e]);
            rny[a][e] = fma4(-ta, sv, ry[a][e]);
            if (!REG) {
                const float se =
                    FIRST ? ew[a][e] : __builtin_fmaf(tc, eq[a][e] - er[a][e], ew[a][e]);
                rne[a][e] = __builtin_fmaf(-ta, se, er[a][e]);
            } else {
                rne[a][e] = 0.0f;
            }
        }
    }

    // ---- every K-th launch: d_(it+1-K) .. d_it and p, finished and stored
    // at once.  The ring's residuals, d and p are loaded only now, after the
    // halo registers above are consumed, and held by nothing below, so that
    // the launch keeps two waves per SIMD.
    if (PH == K - 1) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int a = 0; a < ZU; ++a) {
            const int k = c.k0 + a;
            const int oc = ((REG ? k : min(k, g.nz - 1)) * g.sxy + c.off) >> 2;
            float4 dd = ld4<ST>(d, oc), pv = ld4<ST>(p, oc);
            cgs4_pending<K, PH, K - 1, ST>(ring, it, oc, fa, fb, dd, pv);
            dd = fma4(tb, dd, rz[a + 1]);  // d_it
            pv = fma4(ta, dd, pv);
            if (REG || (c.valid && k < g.nz)) {
                o_d.put(oc, dd);
                o_p.put(oc, pv);
            }
        }
    }

    // ---- w' = A r' and the partials of r'.r', r'.w' ----------------------
    const int srcL = (c.lane + 63) & 63, srcR = (c.lane + 1) & 63;
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        const float sl = lane_value(rn[a + 1].w, srcL), sr = lane_value(rn[a + 1].x, srcR);
        const float xl = !REG && c.needL ? rne[a][0] : sl;
        const float xrv = !REG && c.needR ? rne[a][1] : sr;
        if (REG || (c.valid && k < g.nz)) {
            float q[4];
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
    }
    block_sum<2, false>(acc, lds_out);
    if (threadIdx.x == 0) {
        pg_out[blockIdx.x] = acc[0];
        pd_out[blockIdx.x] = acc[1];
    }
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

template <int NQ, int ST, bool REG, int... PH>
inline const void *cgs4_phase(int ph, std::integer_sequence<int, PH...>) {
    const void *fns[] = {(const void *)k_cg_cgs4<4, false, NQ, ST, kCgsCycle, PH, REG>...};
    return fns[ph];
}

template <int NQ, int ST, bool REG>
inline const void *cgs4_pick(bool first, int ph) {
    if (first) return (const void *)k_cg_cgs4<4, true, NQ, ST, kCgsCycle, 0, REG>;
    return cgs4_phase<NQ, ST, REG>(ph, std::make_integer_sequence<int, kCgsCycle>{});
}

// The k_cg_cgs4 instantiation for launch phase ph (= it mod kCgsCycle).
// wide: more than 1024 partial sums per launch (16 per thread); wt:
// write-through stores and 32-bit load offsets (arrays under 2 GiB), else
// plain stores; reg: a regular tiling (vtiling_regular, write-through only).
inline const void *cgs4_kernel(bool first, int ph, bool wide, bool wt, bool reg) {
    if (wt && reg)
        return wide ? cgs4_pick<16, kStWT, true>(first, ph) : cgs4_pick<4, kStWT, true>(first, ph);
    if (wide)
        return wt ? cgs4_pick<16, kStWT, false>(first, ph)
                  : cgs4_pick<16, kStPlain, false>(first, ph);
    return wt ? cgs4_pick<4, kStWT, false>(first, ph) : cgs4_pick<4, kStPlain, false>(first, ph);
}

inline void cgs4_tail_launch(unsigned nb, hipStream_t st, const CgState *cg, int last,
                             const CgRing &pend, const float4 *d, float4 *p, long n4) {
    k_cg_cgs4_tail<kCgsCycle><<<nb, kThreads, 0, st>>>(cg, last, pend, d, p, n4);
}

}  // namespace
}  // namespace aether
