// k_deform_active.hip -- the adjoint Jacobi step over the voxels that can move.
//
// A step moves voxel j only when it is active, -clip <= psi_j <= Max - eps with psi_j the live
// SDF at j + u_j (KillingFusion.cpp:272-274, 317-319: the update lands only on voxels that
// pass the test).  With u_j = 0 the trilinear read has zero weights and returns phi_live[j]
// exactly, so a voxel with u_j = 0 and phi_live[j] outside the band is inactive, keeps
// u_j = 0 and, by induction, never moves.  The voxels that can move during one deform call
// are therefore fixed when it starts:
//     C = { j : u_j != 0 (any component, -0 counted as 0) or -clip <= phi_live[j] <= Max - eps },
// which is the dense step's own test for running a voxel's stencil (k_adj_flat).
//  * k_active_build streams u and phi_live once, lists C (ascending within a block) and copies
//    u into the field's second buffer, so both ping-pong buffers agree outside C.
//  * k_adj_list runs each listed voxel's stencil (adj_job, the dense step's arithmetic) and
//    writes its new value -- moved or not -- into the output buffer.  Outside C both buffers
//    keep the same values, so every iteration's field equals the dense step's bit for bit.
// C is a shell a few voxels thick around the live surface (a few tenths of a percent of a
// 256^3 grid), so an iteration touches megabytes that stay in the caches instead of streaming
// the whole field through HBM.
#include "kf_adjoint.hpp"

namespace {

constexpr int LB = 64;  // threads per block of the list step: one wavefront, many blocks per CU

template <class S>
__global__ __launch_bounds__(BLOCK) void k_active_build(PhiRead<S> L, const float *__restrict__ u,
                                                        float *__restrict__ u_copy, float lo, float hi,
                                                        unsigned int *__restrict__ list, DeformState *st)
{
    const unsigned int n32 = nvox32(L.g);
    const unsigned int i = blockIdx.x * (unsigned int)BLOCK + threadIdx.x;
    bool cand = false;
    if (i < n32) {
        const float u0 = kf_load_stream(u + i), u1 = kf_load_stream(u + n32 + i), u2 = kf_load_stream(u + 2u * n32 + i);
        const float live = S::load(L.p, (int64_t)i, L.scale);
        u_copy[i] = u0;
        u_copy[n32 + i] = u1;
        u_copy[2u * n32 + i] = u2;
        const bool moved = ((__float_as_uint(u0) | __float_as_uint(u1) | __float_as_uint(u2)) & 0x7fffffffu) != 0u;
        cand = moved || !(live > hi || live < lo);
    }
    // Block-local order is voxel order; one atomic per block places the block's run.
    constexpr int NW = BLOCK / 64;
    __shared__ unsigned int s_off[NW];
    __shared__ unsigned int s_base;
    const unsigned long long bal = __ballot(cand);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        s_off[w] = (unsigned int)__popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int k = 0; k < NW; k++) {
            const unsigned int c = s_off[k];
            s_off[k] = t;
            t += c;
        }
        s_base = t ? atomicAdd(&st->listed, t) : 0u;
    }
    __syncthreads();
    if (cand)
        list[s_base + s_off[w] + (unsigned int)__popcll(bal & ((1ull << lane) - 1ull))] = i;
}

template <class S>
__global__ __launch_bounds__(LB) void k_adj_list(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                 float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                 BlockPartial *partials, int parity, const BlockPartial *prev,
                                                 unsigned int prev_nb, const unsigned int *__restrict__ list,
                                                 unsigned int count)
{
    if (!step_gate<KF_MODE_STEP, LB>(st, prev, prev_nb, parity))
        return;
    const GridDims g = L.g;
    // Consecutive runs of the list (neighbouring voxels) go to one XCD.
    const unsigned int t = kf_xcd_tile(blockIdx.x, gridDim.x) * (unsigned int)LB + threadIdx.x;
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    if (t < count) {
        const unsigned int i = list[t];
        const unsigned int nxy = (unsigned int)g.nx * (unsigned int)g.ny;
        const unsigned int z = i / nxy, r = i - z * nxy;
        const unsigned int y = r / (unsigned int)g.nx, x = r - y * (unsigned int)g.nx;
        adj_job<S, KF_MODE_STEP, true>(L, Cn, u_in, u_out, c, &st->div_slot[parity], nullptr, (int)x, (int)y, (int)z,
                                       mx, esum, act);
    }
    kf_block_publish<LB>(partials, blockIdx.x, esum, mx, act);
}

}  // namespace

bool kf_active_supported(const kf_volume *live) { return 3 * live->n < ((int64_t)1 << 30); }

kf_status kf_launch_active_build(const kf_volume *live, const float *u, float *u_copy, const kf_params *p,
                                 unsigned int *list)
{
    kf_context *ctx = live->ctx;
    KF_CHECK_ARG(kf_active_supported(live), "the active-set sweep needs 3 N < 2^30");
    const AdjConsts c = kf_adj_consts(p);
    const unsigned int nb = (unsigned int)((live->n + BLOCK - 1) / BLOCK);
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_active_build<S>), dim3(nb), dim3(BLOCK), 0, ctx->stream, make_read<S>(live), u, u_copy,
                           c.lo, c.hi, list, ctx->state);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_adj_list(const kf_volume *live, const kf_volume *can, const float *u_in, float *u_out,
                             const kf_params *p, int iter, const unsigned int *list, unsigned int count,
                             unsigned int *nb_out)
{
    kf_context *ctx = live->ctx;
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    const AdjConsts c = kf_adj_consts(p);
    // At least one block: it folds the previous launch and arms the next one.
    const unsigned int nb = count ? (count + LB - 1) / LB : 1u;
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    if (nb_out)
        *nb_out = nb;
    const int parity = iter & 1;
    const BlockPartial *prev = kf_partials(ctx, parity ^ 1);
    const unsigned int prev_nb = iter > 0 ? nb : 0u;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_adj_list<S>), dim3(nb), dim3(LB), 0, ctx->stream, make_read<S>(live), make_read<S>(can),
                           u_in, u_out, c, ctx->state, kf_partials(ctx, parity), parity, prev, prev_nb, list, count);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}
