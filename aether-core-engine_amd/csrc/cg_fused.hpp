// cg_fused.hpp -- one launch per CG iteration, in the single-reduction form of
// CG (A. T. Chronopoulos and C. W. Gear, "s-step iterative methods for
// symmetric linear systems", J. Comput. Appl. Math. 25 (1989) 153-168).
//
// The reference's CG loop (src/fluid.rs:463-487) needs two global reductions
// per iteration (d.Ad, then r.r), so a streaming implementation needs two
// launches.  The single-reduction form reduces gamma = r.r and delta = r.Ar
// together and derives the same scalars (equal in exact arithmetic):
//     beta_i  = gamma_i / gamma_{i-1}
//     mu_i    = delta_i - beta_i gamma_i / alpha_{i-1}      (= d_i . A d_i)
//     alpha_i = gamma_i / mu_i
// and iteration i is
//     d_i = r_i + beta_i d_{i-1};   s_i = A d_i;   p += alpha_i d_i;
//     r_{i+1} = r_i - alpha_i s_i;  gamma_{i+1}, delta_{i+1} from r_{i+1}, A r_{i+1}.
// Only p, r and d are stored: s_i and A r_{i+1} are recomputed from r_i and
// d_{i-1} on a radius-2 neighbourhood.  One launch reads p, r_i, d_{i-1} and
// writes p, r_{i+1}, d_i: 6 array passes, against 8 for the two-launch form.
// Stopping rules are the reference's: gamma_0 <= 1e-24 skips the solve,
// |mu_i| < 1e-28 stops before the update (mu_i is d.q of src/fluid.rs:467),
// gamma < 1e-20 or gamma <= tol^2 |b|^2 stops after it.
//
// Tiling: a workgroup owns a 64 x 4 column tile and marches KZ planes along
// z.  Per plane it loads r_i and d_{i-1} on the tile grown by 2 in x and y
// (68 x 8 cells), forms d_i into an LDS ring, applies A to d_i on the tile
// grown by 1 (66 x 6) to get r_{i+1} into a second ring, then applies A to
// r_{i+1} on the tile itself and writes the outputs.  Global loads run PF
// planes ahead of their use; the plane loop is fully unrolled so that every
// ring slot is a compile-time index.
#pragma once
#include "fluid_kernels.hpp"

namespace aether {
namespace {

constexpr int kFX = 64, kFY = 4;                                  // output tile
constexpr int kR2X = kFX + 4, kR2Y = kFY + 4, kR2 = kR2X * kR2Y;  // 68 x 8 = 544
constexpr int kR1X = kFX + 2, kR1Y = kFY + 2, kR1 = kR1X * kR1Y;  // 66 x 6 = 396
constexpr int kR2Per = (kR2 + kThreads - 1) / kThreads;           // 3
constexpr int kR1Per = (kR1 + kThreads - 1) / kThreads;           // 2
static_assert(kFX == kBX && kFY == kBY && kFX * kFY == kThreads, "tile = workgroup shape");
constexpr int kFusedKZ = 16;  // z planes per workgroup
constexpr int kFusedPF = 4;   // planes of loads in flight ahead of their use

// delta_0 = r_0 . A r_0 partials over the initialisation tiling (the other
// half of the first iteration's reduction; gamma_0 comes with r_0).  When
// w_out is set, A r_0 is also stored there (0 off the interior).
template <typename T, int ZU>
__device__ __forceinline__ void body_cg_rar(const Col &c, int slot, const Dims &g,
                                            const T *__restrict__ r, double *__restrict__ part,
                                            T *__restrict__ w_out) {
    __shared__ double lds[8];
    T rz[ZU + 2], rn[ZU][4];
#pragma unroll
    for (int a = 0; a < ZU + 2; ++a) rz[a] = r[col_id(g, c, c.k0 - 1 + a)];
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int id = col_id(g, c, c.k0 + a);
        rn[a][0] = r[id + c.xm];
        rn[a][1] = r[id + c.xp];
        rn[a][2] = r[id + c.ym];
        rn[a][3] = r[id + c.yp];
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int a = 0; a < ZU; ++a) {
        const int k = c.k0 + a;
        if (!c.valid || k >= g.nz) continue;
        T q = T(0);
        if (is_interior(g, c.i, c.j, k)) {
            q = lap7<T>(rz[a + 1], rn[a][0], rn[a][1], rn[a][2], rn[a][3], rz[a], rz[a + 2],
                        c.i > 1, c.i + 2 < g.nx, c.j > 1, c.j + 2 < g.ny, k > 1, k + 2 < g.nz);
            acc[0] += (double)rz[a + 1] * (double)q;
        }
        if (w_out) w_out[col_id(g, c, k)] = q;
    }
    block_sum<1>(acc, lds);
    if (tid_lin() == 0) part[slot] = acc[0];
}

template <typename T, int ZU>
__global__ __launch_bounds__(kThreads) void k_cg_rar(Dims g, Tiling tl, const T *__restrict__ r,
                                                     double *__restrict__ part,
                                                     T *__restrict__ w_out) {
    body_cg_rar<T, ZU>(col_of<ZU>(g, tl), blockIdx.x, g, r, part, w_out);
}

// Iteration `it` of the single-reduction CG.  pg_in/pd_in: partials of
// gamma_it and delta_it (nb_in of them); pg_out/pd_out: this launch's
// partials of gamma_{it+1}, delta_{it+1}.  FIRST: it == 0 (reads |b|^2
// partials, no d_{-1}).  PF: planes of global loads in flight ahead of use.
//
// Step s of the plane loop runs three stages and then one barrier:
//   C: A r_{i+1}, reductions and outputs on plane k0 - 5 + s
//   B: r_{i+1} = r_i - alpha A d_i on plane k0 - 3 + s (66 x 6 region)
//   A: d_i, r_i of plane k0 - 1 + s into the rings (68 x 8 region)
// C of step s reads what B wrote in steps s-1 .. s-3 and B reads what A
// wrote in steps s-1 .. s-3, all behind at least one barrier; the ring
// depths (d: 5, r_i: 3, r_{i+1}: 4 planes) keep every slot written in step
// s apart from every slot read in step s.
template <typename T, int KZ, int PF, bool FIRST>
__global__ __launch_bounds__(kThreads) void k_cg_fused(
    Dims g, Tiling tl, int it, double rel_tol2, int nb_in, const double *__restrict__ pg_in,
    const double *__restrict__ pd_in, const double *__restrict__ pbb, double *__restrict__ pg_out,
    double *__restrict__ pd_out, CgState *st, const T *__restrict__ r_in, T *__restrict__ r_out,
    const T *__restrict__ d_in, T *__restrict__ d_out, T *__restrict__ p, int *done_mirror,
    int solve_id) {
    __shared__ T sD[5][kR2];  // d_i
    __shared__ T sR[3][kR2];  // r_i
    __shared__ T sP[4][kR1];  // r_{i+1}
    __shared__ double lds[16];
    __shared__ int flag_lds[1];
    constexpr int NL = KZ + 4;  // planes loaded: k0-2 .. k0+KZ+1 (load index l = plane - k0 + 2)
    constexpr int NS = KZ + 6;  // loop steps: A handles load index s, B s-2, C s-4 (<= KZ+1)

    int bx, by, bz;
    tile_coords(tl, bx, by, bz);
    const int x0 = bx * kFX, y0 = by * kFY, k0 = bz * KZ;
    const int t = tid_lin();
    const int tx = threadIdx.x, ty = threadIdx.y;

    // cells of the 68 x 8 load region handled by this thread
    int offA[kR2Per];
    bool okA[kR2Per];
#pragma unroll
    for (int q = 0; q < kR2Per; ++q) {
        const int c = t + q * kThreads;
        const int cy = c / kR2X, cx = c - cy * kR2X;
        const int gx = x0 - 2 + cx, gy = y0 - 2 + cy;
        okA[q] = c < kR2 && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
        offA[q] = min(max(gy, 0), g.ny - 1) * g.nx + min(max(gx, 0), g.nx - 1);
    }
    const bool third = t < kR2 - (kR2Per - 1) * kThreads;  // only wave 0's first 32 lanes
    const int offT = min(y0 + ty, g.ny - 1) * g.nx + min(x0 + tx, g.nx - 1);

    // Raw prefetched values, by load index; masks are applied where they are
    // consumed so that no instruction right after a load waits for it.
    T vr[NL][kR2Per], vd[NL][kR2Per], vp[NL];
    auto issue = [&](int l) {
        const int L = min(max(k0 - 2 + l, 0), g.nz - 1);
        const int base = L * g.sxy;
#pragma unroll
        for (int q = 0; q < kR2Per; ++q) {
            if (q == kR2Per - 1 && !third) continue;
            vr[l][q] = r_in[base + offA[q]];
            if (!FIRST) vd[l][q] = d_in[base + offA[q]];
        }
        if (l >= 2 && l < KZ + 2) vp[l] = p[base + offT];
    };

    const int flag_reg = st->done;
    EarlyPartials eg, ed, eb;
    eg.load(pg_in, nb_in);
    ed.load(pd_in, nb_in);
    if (FIRST) eb.load(pbb, nb_in);
    double gprev = 1.0, aprev = 1.0, b2 = 0.0;
    if (!FIRST) {
        gprev = st->cg_gamma[(it - 1) & 1];
        aprev = st->cg_alpha[(it - 1) & 1];
        b2 = st->b2;
    }
#pragma unroll
    for (int l = 0; l < PF && l < NL; ++l) issue(l);

    // scalars (every workgroup reaches the same decision)
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
    const T tb = (T)beta, ta = (T)alpha;

    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (s + PF < NL) issue(s + PF);
        // C: A r_{i+1}, reductions and outputs on the plane of load index l = s - 4
        if (s >= 4 && s - 4 >= 2 && s - 4 < KZ + 2) {
            const int l = s - 4;
            const int n = k0 - 2 + l;
            const int gx = x0 + tx, gy = y0 + ty;
            const int c1 = (ty + 1) * kR1X + tx + 1;
            const T rc = sP[l & 3][c1];
            T w = T(0);
            if (is_interior(g, gx, gy, n))
                w = lap7<T>(rc, sP[l & 3][c1 - 1], sP[l & 3][c1 + 1], sP[l & 3][c1 - kR1X],
                            sP[l & 3][c1 + kR1X], sP[(l - 1) & 3][c1], sP[(l + 1) & 3][c1],
                            gx > 1, gx + 2 < g.nx, gy > 1, gy + 2 < g.ny, n > 1, n + 2 < g.nz);
            acc[0] += (double)rc * (double)rc;
            acc[1] += (double)rc * (double)w;
            if (gx < g.nx && gy < g.ny && n < g.nz) {
                const int id = n * g.sxy + gy * g.nx + gx;
                const T dv = sD[l % 5][(ty + 2) * kR2X + tx + 2];
                r_out[id] = rc;
                d_out[id] = dv;
                p[id] = vp[l] + ta * dv;
            }
        }
        // B: r_{i+1} on the 66 x 6 region of the plane of load index l = s - 2
        if (s >= 2 && s - 2 >= 1 && s - 2 < NL - 1) {
            const int l = s - 2;
            const int m = k0 - 2 + l;
#pragma unroll
            for (int q = 0; q < kR1Per; ++q) {
                const int c1 = t + q * kThreads;
                if (c1 < kR1) {
                    const int cy = c1 / kR1X, cx = c1 - cy * kR1X;
                    const int gx = x0 - 1 + cx, gy = y0 - 1 + cy;
                    const int c2 = (cy + 1) * kR2X + cx + 1;
                    T sv = T(0);
                    if (is_interior(g, gx, gy, m))
                        sv = lap7<T>(sD[l % 5][c2], sD[l % 5][c2 - 1], sD[l % 5][c2 + 1],
                                     sD[l % 5][c2 - kR2X], sD[l % 5][c2 + kR2X],
                                     sD[(l + 4) % 5][c2], sD[(l + 1) % 5][c2], gx > 1,
                                     gx + 2 < g.nx, gy > 1, gy + 2 < g.ny, m > 1, m + 2 < g.nz);
                    sP[l & 3][c1] = sR[l % 3][c2] - ta * sv;
                }
            }
        }
        // A: d_i and r_i of the plane of load index l = s
        if (s < NL) {
            const int l = s;
            const int L = k0 - 2 + l;
            const bool okz = L >= 0 && L < g.nz;
#pragma unroll
            for (int q = 0; q < kR2Per; ++q) {
                if (q == kR2Per - 1 && !third) continue;
                const int c = t + q * kThreads;
                const bool ok = okz && okA[q];
                const T rv = ok ? vr[l][q] : T(0);
                T dv = rv;
                if (!FIRST) dv = rv + tb * (ok ? vd[l][q] : T(0));
                sD[l % 5][c] = dv;
                sR[l % 3][c] = rv;
            }
        }
        __syncthreads();
    }
    block_sum<2>(acc, lds);
    if (t == 0) {
        pg_out[blockIdx.x] = acc[0];
        pd_out[blockIdx.x] = acc[1];
    }
}

}  // namespace
}  // namespace aether
