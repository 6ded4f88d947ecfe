// cg_resident.hpp -- the whole CG solve (src/fluid.rs:445-493) as one
// cooperative launch whose vectors stay in registers.
//
// Each workgroup owns a 64 x 4 x ZT tile; each thread owns one z column of
// ZT cells and keeps p, r, d and q for them in registers for the whole
// solve.  Per iteration only the tile halos travel through memory: a thread
// publishes r and d of its cells, and reads r and d of its x/y neighbours
// (and of the cells just past its column ends) to rebuild d = r + beta*d
// there.  Two grid-wide barriers per iteration carry the two reductions of
// the reference's loop (d.q, then r.r); every workgroup re-reduces the
// partials in the same fixed order, so all take identical decisions.
//
// The barrier is bounded: a workgroup that waits longer than kSpinLimit
// polls sets GridSync::error and leaves; the others then time out at their
// next barrier, so the launch always terminates and the host reports the
// error instead of hanging.  The launch is cooperative (the runtime refuses
// it when the tiles cannot all be resident), so in normal operation every
// workgroup reaches every barrier.
#pragma once
#include <hip/hip_cooperative_groups.h>

#include "fluid_kernels.hpp"

namespace aether {

// Barrier words, each on its own 2 KiB line so that the arrival atomics and
// the pollers do not queue behind each other in one memory channel.
struct GridSync {
    unsigned int count;  // arrivals at the current barrier
    unsigned int pad0[511];
    unsigned int gen;  // barrier generation
    unsigned int pad1[511];
    int error;  // set when a barrier wait exceeded the spin limit
    int pad2[511];
};

constexpr long kSpinLimit = 1L << 20;

// Grid-wide barrier.  BMODE 0: counter + generation word with agent-scope
// release/acquire fences, bounded wait.  BMODE 1: the runtime's cooperative
// grid barrier (cooperative_groups grid sync), bracketed by the same fences.
// `gen` is the generation this workgroup waits to leave (thread 0 only).
// Returns false on timeout (for every thread of the block).
template <int BMODE>
__device__ __forceinline__ bool grid_barrier(GridSync *gs, unsigned nblocks, unsigned &gen,
                                             int *fail_lds) {
    if constexpr (BMODE == 1) {
        __threadfence();
        cooperative_groups::this_grid().sync();
        __threadfence();
        return true;
    } else {
        __syncthreads();
        if (tid_lin() == 0) {
            int fail = 0;
            __threadfence();  // release: this workgroup's global writes
            const unsigned arrived = atomicAdd(&gs->count, 1u);
            if (arrived == nblocks - 1) {
                atomicExch(&gs->count, 0u);
                __threadfence();
                atomicAdd(&gs->gen, 1u);
            } else {
                long spins = 0;
                while (__hip_atomic_load(&gs->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                       gen) {
                    __builtin_amdgcn_s_sleep(4);
                    if (++spins > kSpinLimit) {
                        fail = 1;
                        atomicExch(&gs->error, 1);
                        break;
                    }
                }
            }
            __threadfence();  // acquire: other workgroups' writes
            gen += 1;
            fail_lds[0] = fail;
        }
        __syncthreads();
        return fail_lds[0] == 0;
    }
}

// Fixed-order sum of nb partials (any nb), identical in every workgroup.
__device__ __forceinline__ double sum_all(const double *__restrict__ part, int nb, double *lds) {
    double acc[1] = {0.0};
    for (int a = tid_lin(); a < nb; a += kThreads) acc[0] += part[a];
    block_sum<1>(acc, lds);
    return acc[0];
}

// Buffer-resource access with 32-bit byte offsets: one VGPR per address and
// hardware range checking (an offset outside [0, bytes) reads 0 and writes
// nothing), which also covers the halo reads that step past the grid edges.
// Every offset is passed in the VGPR operand (soffset stays 0), so the range
// check sees the full offset.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t rs, int voff) {
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, voff, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, voff, 0, 0);
}

// Minimum waves per SIMD that the tile size needs for all tiles of a 128^3
// grid to be resident at once (ZT = 8: 4 workgroups per CU; ZT = 16: 2).
template <int ZT>
struct ResidentWaves {
    static constexpr int value = ZT <= 8 ? 4 : 2;
};

template <typename T, int ZT, int BMODE>
__global__ __launch_bounds__(kThreads, ResidentWaves<ZT>::value) void k_cg_resident(
    Dims g, Tiling tl, int max_iters, double rel_tol2, const double *__restrict__ part_rr0,
    const double *__restrict__ part_bb0, int nb0, double *__restrict__ part_dq,
    double *__restrict__ part_rr, T *__restrict__ pg, T *__restrict__ rg, T *__restrict__ dg,
    CgState *st, int *__restrict__ iters_log, GridSync *gs) {
    __shared__ double lds[8];
    __shared__ int fail_lds[1];
    unsigned gen = 0;
    if (tid_lin() == 0) gen = __hip_atomic_load(&gs->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int bx, by, bz;
    tile_coords(tl, bx, by, bz);
    const int i = bx * kBX + threadIdx.x, j = by * kBY + threadIdx.y, k0 = bz * ZT;
    const bool col = i < g.nx && j < g.ny;
    const unsigned nblocks = gridDim.x;
    const bool lead = blockIdx.x == 0 && tid_lin() == 0;
    constexpr int S = (int)sizeof(T);
    const int bytes = g.sxy * g.nz * S;
    const int rowb = g.nx * S, planeb = g.sxy * S;
    const auto rsP = buf_rsrc(pg, bytes), rsR = buf_rsrc(rg, bytes), rsD = buf_rsrc(dg, bytes);
    const int off0 = ((k0 * g.ny + j) * g.nx + i) * S;  // byte offset of the column's first cell

    T p[ZT], r[ZT], d[ZT], q[ZT];
#pragma unroll
    for (int a = 0; a < ZT; ++a) {
        p[a] = bload<T>(rsP, off0 + a * planeb, 0);
        r[a] = bload<T>(rsR, off0 + a * planeb, 0);
        d[a] = T(0);
        q[a] = T(0);
    }
    auto live = [&](int a) { return col && k0 + a < g.nz; };
    auto inner = [&](int a) { return live(a) && is_interior(g, i, j, k0 + a); };

    const double delta0 = sum_all(part_rr0, nb0, lds);
    const double b2 = sum_all(part_bb0, nb0, lds);
    double delta = delta0;
    int iters = 0, converged = 0;
    bool stop = !(delta0 > 1e-24) || (rel_tol2 > 0.0 && delta0 <= rel_tol2 * b2);
    if (stop) converged = 1;
    double beta = 0.0;
    bool ok = true;

    for (int it = 0; it < max_iters && !stop; ++it) {
        // ---- phase A: d = r + beta d, q = A d, partial d.q -------------
        const T tb = (T)beta;
        const bool first = it == 0;
        // neighbour values of d rebuilt from published r and d_old
        auto dnb = [&](int voff) -> T {
            const T rv = bload<T>(rsR, voff, 0);
            return first ? rv : rv + tb * bload<T>(rsD, voff, 0);
        };
#pragma unroll
        for (int a = 0; a < ZT; ++a) d[a] = first ? r[a] : r[a] + tb * d[a];
        const T below = dnb(off0 - planeb);
        const T above = dnb(off0 + ZT * planeb);
        double acc[1] = {0.0};
#pragma unroll
        for (int a = 0; a < ZT; ++a) {
            const int oa = off0 + a * planeb;
            const T xm = dnb(oa - S), xp = dnb(oa + S);
            const T ym = dnb(oa - rowb), yp = dnb(oa + rowb);
            const T zm = a == 0 ? below : d[a - 1];
            const T zp = a == ZT - 1 ? above : d[a + 1];
            const int k = k0 + a;
            q[a] = inner(a) ? lap7<T>(d[a], xm, xp, ym, yp, zm, zp, i > 1, i + 2 < g.nx, j > 1,
                                      j + 2 < g.ny, k > 1, k + 2 < g.nz)
                            : T(0);
            if (inner(a)) acc[0] += (double)d[a] * (double)q[a];
        }
        block_sum<1>(acc, lds);
        if (tid_lin() == 0) part_dq[blockIdx.x] = acc[0];
        ok = grid_barrier<BMODE>(gs, nblocks, gen, fail_lds);
        if (!ok) break;

        // ---- phase B: alpha, p += alpha d, r -= alpha q, partial r.r ---
        const double dq = sum_all(part_dq, (int)nblocks, lds);
        if (fabs(dq) < 1e-28) {
            stop = true;
            converged = 1;
            break;
        }
        const T ta = (T)(delta / dq);
        double acc2[1] = {0.0};
#pragma unroll
        for (int a = 0; a < ZT; ++a) {
            p[a] = p[a] + ta * d[a];
            r[a] = r[a] - ta * q[a];
            if (live(a)) {
                acc2[0] += (double)r[a] * (double)r[a];
                bstore<T>(r[a], rsR, off0 + a * planeb);
                bstore<T>(d[a], rsD, off0 + a * planeb);
            }
        }
        block_sum<1>(acc2, lds);
        if (tid_lin() == 0) part_rr[blockIdx.x] = acc2[0];
        ok = grid_barrier<BMODE>(gs, nblocks, gen, fail_lds);
        if (!ok) break;
        const double rr = sum_all(part_rr, (int)nblocks, lds);
        iters = it + 1;
        if (rr < 1e-20 || (rel_tol2 > 0.0 && rr <= rel_tol2 * b2)) {
            delta = rr;
            stop = true;
            converged = 1;
            break;
        }
        beta = rr / delta;
        delta = rr;
    }

#pragma unroll
    for (int a = 0; a < ZT; ++a)
        if (live(a)) bstore<T>(p[a], rsP, off0 + a * planeb);
    if (lead) {
        st->b2 = b2;
        st->r2_init = delta0;
        st->r2_final = delta;
        st->iters = iters;
        st->converged = ok ? converged : -1;
        st->done = 1;
        if (iters_log) *iters_log = iters;
    }
}

__global__ void k_grid_sync_reset(GridSync *gs) {
    if (threadIdx.x == 0) {
        gs->count = 0;
        gs->error = 0;
    }
}

}  // namespace aether
