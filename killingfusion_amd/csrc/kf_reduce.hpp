// kf_reduce.hpp -- wavefront/block reductions of a deform iteration's FP64 max |du|,
// FP64 energy sum (fixed order) and active count.
#pragma once

#include "kf_internal.hpp"

#include <climits>

// Wavefront (64 lanes) butterfly reductions.
__device__ inline double kf_wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double kf_wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline unsigned long long kf_wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide reduction of (sum, max, count); result valid in thread 0.  Uses a
// fixed combination order, so the FP64 sum is reproducible for a fixed launch shape.  A
// wavefront with wave_has false (uniform over it) holds (+0, +0, 0) in every lane and skips its
// reductions, whose result would be the same zeros (dense adjoint step, 256^3 BF16: 41.6-41.8
// against 41.5-42.3 us).
template <int BLOCK>
__device__ inline void kf_block_reduce3(double &sum, double &mx, unsigned long long &cnt, bool wave_has = true)
{
    constexpr int NW = BLOCK / 64;
    __shared__ double s_sum[NW];
    __shared__ double s_max[NW];
    __shared__ unsigned long long s_cnt[NW];
    if (wave_has) {
        sum = kf_wave_sum(sum);
        mx = kf_wave_max(mx);
        cnt = kf_wave_sum_u64(cnt);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s_sum[wid] = sum;
        s_max[wid] = mx;
        s_cnt[wid] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, m = 0.0;
        unsigned long long c = 0;
        for (int w = 0; w < NW; w++) {
            a += s_sum[w];
            m = fmax(m, s_max[w]);
            c += s_cnt[w];
        }
        sum = a;
        mx = m;
        cnt = c;
    }
}

// Thread 0 stores the block's (sum, max, count) in its slot; k_finalize (kf_api.hip) folds
// the slots in index order after the launch.  No same-address atomics: with tens of
// thousands of blocks those serialise in L2 and cost more than the stencil itself.
template <int BLOCK>
__device__ inline void kf_block_publish(BlockPartial *parts, unsigned int slot, double sum, double mx,
                                        unsigned long long cnt, bool wave_has = true)
{
    kf_block_reduce3<BLOCK>(sum, mx, cnt, wave_has);
    if (threadIdx.x == 0) {
        BlockPartial p;
        p.sum = sum;
        p.max = mx;
        p.cnt = cnt;
        p.pad = 0ull;
        parts[slot] = p;
    }
}

// One whole block folds the nb partials of one step launch (slot order fixed: strided by
// BLOCK, eight loads in flight per thread) into the iteration state.  `parity` names the
// launch's divergence slot.  A launch folded after the state is done changes nothing.
// Uses the LDS of kf_block_reduce3: a caller that reduces again afterwards puts a barrier
// in between.
template <int BLOCK>
__device__ inline void kf_fold_launch(DeformState *st, const BlockPartial *parts, unsigned int nb, int parity)
{
    double s = 0.0, m = 0.0;
    unsigned long long c = 0;
    for (unsigned int b0 = threadIdx.x; b0 < nb; b0 += 8 * BLOCK) {
        BlockPartial p[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const unsigned int b = b0 + (unsigned int)j * BLOCK;
            if (b < nb)
                p[j] = parts[b];
            else
                p[j] = BlockPartial{0.0, 0.0, 0ull, 0ull};
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            s += p[j].sum;
            m = fmax(m, p[j].max);
            c += p[j].cnt;
        }
    }
    kf_block_reduce3<BLOCK>(s, m, c);
    if (threadIdx.x == 0 && !st->done) {
        const long long dv = st->div_slot[parity & 1];
        st->div_slot[parity & 1] = LLONG_MAX;
        st->last_max = m;
        st->energy = s;
        st->last_active = c;
        st->iters += 1;
        if (dv != LLONG_MAX)
            st->diverged = dv;
        if (dv != LLONG_MAX || (st->stop > 0.0 && m < st->stop))
            st->done = 1;
    }
}
