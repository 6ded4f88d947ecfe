// cg_fused4.hpp -- one launch per CG iteration (the single-reduction form of
// cg_fused.hpp) with four x-cells per thread and whole rows per workgroup.
//
// Iteration i reads p, r_i, d_{i-1} and writes p, r_{i+1}, d_i; s_i = A d_i
// and w = A r_{i+1} are recomputed on a radius-2 neighbourhood.  A workgroup
// owns whole x-rows (txv = nx/4 float4 lanes, so the x-edges are walls) of ty
// consecutive rows and ZU planes; each thread owns one float4 x ZU column.
//   phase 1: loads: own column r, d_{i-1} on planes k0-2 .. k0+ZU+1, p on the
//            ZU own planes; the halo rows j0-2, j0-1, j0+ty, j0+ty+1 are
//            shared out among the threads of each lane column.  d_i = r +
//            beta d_{i-1} goes to LDS (rows j0-2 .. j0+ty+1, planes k0-1 ..
//            k0+ZU).
//   phase 2: s = A d_i and r_{i+1} = r_i - alpha s on the own column (planes
//            k0-1 .. k0+ZU) and on the halo rows j0-1, j0+ty (own planes);
//            r_{i+1} on the own planes goes to LDS (rows j0-1 .. j0+ty).
//   phase 3: w = A r_{i+1}, the gamma/delta partials and the outputs.
// y and x neighbours come from LDS, z neighbours from the thread's own
// registers.  Needs nx % 4 == 0 and nx <= 256.
#pragma once
#include "cg_vec.hpp"

namespace aether {

template <int ZU>
struct Fused4Geom {
    static constexpr int NA = ZU + 2;           // LDS planes of d_i: k0-1 .. k0+ZU
    static constexpr int NH = 4 * ZU + 4;       // halo cells per lane column
    static constexpr int MAXH = (NH + 3) / 4;   // per thread (ty >= 4)
    static constexpr int DCAP = 512;            // (ty + 4) * txv <= 256 + 4 * 64
    static constexpr int RCAP = 384;            // (ty + 2) * txv <= 256 + 2 * 64
};

template <int ZU, bool FIRST, int NQ>
__global__ __launch_bounds__(kThreads) void k_cg_fused4(
    Dims g, VTiling vt, int it, double rel_tol2, int nb_in, const double *__restrict__ pg_in,
    const double *__restrict__ pd_in, const double *__restrict__ pbb, double *__restrict__ pg_out,
    double *__restrict__ pd_out, CgState *st, const float *__restrict__ r_in,
    float *__restrict__ r_out, const float *__restrict__ d_in, float *__restrict__ d_out,
    float *__restrict__ p, int *done_mirror, int solve_id) {
    using G = Fused4Geom<ZU>;
    __shared__ float4 sD[G::NA * G::DCAP];
    __shared__ float4 sR[ZU * G::RCAP];
    __shared__ double lds[16];
    __shared__ int flag_lds[1];

    int bx, by, bz;
    vtile_coords(vt, bx, by, bz);
    const int TX = vt.txv, TY = vt.ty;
    const int t = threadIdx.x;
    const int txv = t % TX, ty = t / TX;
    const bool member = ty < TY;  // threads past TX * TY idle
    const int i0 = 4 * txv, j0 = by * TY, k0 = bz * ZU;
    const int j = j0 + ty;
    const bool valid = member && j < g.ny;
    const int txm = max(txv - 1, 0), txp = min(txv + 1, TX - 1);
    // LDS addressing: row R relative to j0, plane a (plane k0 - 1 + a)
    auto dix = [&](int a, int R, int x) { return (a * (TY + 4) + R + 2) * TX + x; };
    auto rix = [&](int a, int R, int x) { return ((a - 1) * (TY + 2) + R + 1) * TX + x; };
    auto goff = [&](int kk, int jj) {
        return min(max(kk, 0), g.nz - 1) * g.sxy + min(max(jj, 0), g.ny - 1) * g.nx + i0;
    };
    const float4 *r4 = reinterpret_cast<const float4 *>(r_in);
    const float4 *d4 = reinterpret_cast<const float4 *>(d_in);
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);

    // ---- phase 1 loads (all issued before the scalars are reduced) -------
    const int flag_reg = st->done;
    EarlyPartialsT<NQ> eg, ed, eb;
    eg.load(pg_in, nb_in);
    ed.load(pd_in, nb_in);
    if (FIRST) eb.load(pbb, nb_in);
    double gprev = 1.0, aprev = 1.0, b2 = 0.0;
    if (!FIRST) {
        gprev = st->cg_gamma[(it - 1) & 1];
        aprev = st->cg_alpha[(it - 1) & 1];
        b2 = st->b2;
    }
    float4 rc[ZU + 4], dc[ZU + 4], pc[ZU];  // own column, planes k0-2 .. k0+ZU+1
#pragma unroll
    for (int c = 0; c < ZU + 4; ++c) {
        const int o = goff(k0 - 2 + c, j) >> 2;
        rc[c] = r4[o];
        dc[c] = FIRST ? z4 : d4[o];
    }
#pragma unroll
    for (int a = 0; a < ZU; ++a) pc[a] = reinterpret_cast<const float4 *>(p)[goff(k0 + a, j) >> 2];
    // halo cells of this lane column: items h = ty, ty + TY, ...; items
    // 0 .. 2NA-1 are rows -1 and TY on planes 0 .. NA-1 (also get s, r'),
    // the rest rows -2 and TY+1 on planes 1 .. ZU
    float4 hr[G::MAXH], hd[G::MAXH];
    int hR[G::MAXH], hA[G::MAXH];
    bool hOn[G::MAXH];
#pragma unroll
    for (int q = 0; q < G::MAXH; ++q) {
        const int h = ty + q * TY;
        hOn[q] = member && h < G::NH;
        int R, a;
        if (h < 2 * G::NA) {
            R = h < G::NA ? -1 : TY;
            a = h % G::NA;
        } else {
            const int hb = h - 2 * G::NA;
            R = hb < ZU ? -2 : TY + 1;
            a = 1 + hb % ZU;
        }
        hR[q] = R;
        hA[q] = a;
        const int o = goff(k0 - 1 + a, j0 + R) >> 2;
        hr[q] = hOn[q] ? r4[o] : z4;
        hd[q] = (hOn[q] && !FIRST) ? d4[o] : z4;
    }

    // ---- scalars (every workgroup reaches the same decision) -------------
    if (t == 0) flag_lds[0] = flag_reg;
    const double gam = eg.total(pg_in, nb_in, lds);  // its barrier publishes flag_lds
    if (flag_lds[0]) return;
    const double del = ed.total(pd_in, nb_in, lds);
    const bool lead = blockIdx.x == 0 && t == 0;
    double beta = 0.0, mu = del;
    bool stop;
    if (FIRST) {
        b2 = eb.total(pbb, nb_in, lds);
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
            st->done = 1;
            st->iters = it;
            st->r2_final = gam;
            st->converged = 1;
            *done_mirror = solve_id;
        }
        return;
    }
    const double alpha = gam / mu;
    if (lead) {
        st->cg_gamma[it & 1] = gam;
        st->cg_alpha[it & 1] = alpha;
    }
    const float tb = (float)beta, ta = (float)alpha;

    // ---- phase 1: d_i into LDS -------------------------------------------
#pragma unroll
    for (int c = 0; c < ZU + 4; ++c) dc[c] = FIRST ? rc[c] : axpy4(rc[c], tb, dc[c]);
    if (member) {
#pragma unroll
        for (int a = 0; a < G::NA; ++a) sD[dix(a, ty, txv)] = dc[a + 1];
    }
#pragma unroll
    for (int q = 0; q < G::MAXH; ++q) {
        hd[q] = FIRST ? hr[q] : axpy4(hr[q], tb, hd[q]);
        if (hOn[q]) sD[dix(hA[q], hR[q], txv)] = hd[q];
    }
    __syncthreads();

    // ---- phase 2: s = A d_i, r' = r - alpha s ------------------------------
    // s on four cells of row R (relative), plane a, with z neighbours zm, zp
    auto sfour = [&](int R, int a, const float4 &ctr, const float4 &zm, const float4 &zp) {
        const float4 xl = sD[dix(a, R, txm)], xr = sD[dix(a, R, txp)];
        const float4 ym = sD[dix(a, R - 1, txv)], yp = sD[dix(a, R + 1, txv)];
        const int jj = j0 + R, kk = k0 - 1 + a;
        float q[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = i0 + e;
            q[e] = 0.0f;
            if (is_interior(g, i, jj, kk))
                q[e] = lap7<float>(f4(ctr, e), e == 0 ? xl.w : f4(ctr, e - 1),
                                   e == 3 ? xr.x : f4(ctr, e + 1), f4(ym, e), f4(yp, e),
                                   f4(zm, e), f4(zp, e), i > 1, i + 2 < g.nx, jj > 1,
                                   jj + 2 < g.ny, kk > 1, kk + 2 < g.nz);
        }
        return make_float4(q[0], q[1], q[2], q[3]);
    };
    float4 rp[G::NA];  // r' on the own column, planes k0-1 .. k0+ZU
    if (member) {
#pragma unroll
        for (int a = 0; a < G::NA; ++a) {
            const float4 s = sfour(ty, a, dc[a + 1], dc[a], dc[a + 2]);
            rp[a] = make_float4(rc[a + 1].x - ta * s.x, rc[a + 1].y - ta * s.y,
                                rc[a + 1].z - ta * s.z, rc[a + 1].w - ta * s.w);
            if (a >= 1 && a <= ZU) sR[rix(a, ty, txv)] = rp[a];
        }
    }
#pragma unroll
    for (int q = 0; q < G::MAXH; ++q) {
        const int a = hA[q];
        // rows -1 and TY on the own planes also need r'
        if (hOn[q] && (hR[q] == -1 || hR[q] == TY) && a >= 1 && a <= ZU) {
            const float4 s = sfour(hR[q], a, hd[q], sD[dix(a - 1, hR[q], txv)],
                                   sD[dix(a + 1, hR[q], txv)]);
            sR[rix(a, hR[q], txv)] = make_float4(hr[q].x - ta * s.x, hr[q].y - ta * s.y,
                                                 hr[q].z - ta * s.z, hr[q].w - ta * s.w);
        }
    }
    __syncthreads();

    // ---- phase 3: w = A r', reductions, outputs ---------------------------
    double acc[2] = {0.0, 0.0};
    if (valid) {
        float4 *p4 = reinterpret_cast<float4 *>(p);
        float4 *ro4 = reinterpret_cast<float4 *>(r_out);
        float4 *do4 = reinterpret_cast<float4 *>(d_out);
#pragma unroll
        for (int a = 1; a <= ZU; ++a) {
            const int kk = k0 - 1 + a;
            if (kk >= g.nz) break;
            const float4 ctr = rp[a];
            const float4 xl = sR[rix(a, ty, txm)], xr = sR[rix(a, ty, txp)];
            const float4 ym = sR[rix(a, ty - 1, txv)], yp = sR[rix(a, ty + 1, txv)];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = i0 + e;
                float w = 0.0f;
                if (is_interior(g, i, j, kk))
                    w = lap7<float>(f4(ctr, e), e == 0 ? xl.w : f4(ctr, e - 1),
                                    e == 3 ? xr.x : f4(ctr, e + 1), f4(ym, e), f4(yp, e),
                                    f4(rp[a - 1], e), f4(rp[a + 1], e), i > 1, i + 2 < g.nx,
                                    j > 1, j + 2 < g.ny, kk > 1, kk + 2 < g.nz);
                acc[0] += (double)f4(ctr, e) * (double)f4(ctr, e);
                acc[1] += (double)f4(ctr, e) * (double)w;
            }
            const int o = (kk * g.sxy + j * g.nx + i0) >> 2;
            st4<false>(ro4 + o, ctr);
            st4<false>(do4 + o, dc[a + 1]);
            st4<false>(p4 + o, axpy4(pc[a - 1], ta, dc[a + 1]));
        }
    }
    block_sum<2>(acc, lds);
    if (t == 0) {
        pg_out[blockIdx.x] = acc[0];
        pd_out[blockIdx.x] = acc[1];
    }
}

}  // namespace aether
