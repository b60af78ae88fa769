// kf_reduce.hpp -- wavefront/block reductions of a deform iteration's FP64 max |du|,
// FP64 energy sum (fixed order) and active count.
#pragma once

#include "kf_internal.hpp"

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
// fixed combination order, so the FP64 sum is reproducible for a fixed launch shape.
template <int BLOCK>
__device__ inline void kf_block_reduce3(double &sum, double &mx, unsigned long long &cnt)
{
    constexpr int NW = BLOCK / 64;
    __shared__ double s_sum[NW];
    __shared__ double s_max[NW];
    __shared__ unsigned long long s_cnt[NW];
    sum = kf_wave_sum(sum);
    mx = kf_wave_max(mx);
    cnt = kf_wave_sum_u64(cnt);
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
                                        unsigned long long cnt)
{
    kf_block_reduce3<BLOCK>(sum, mx, cnt);
    if (threadIdx.x == 0) {
        BlockPartial p;
        p.sum = sum;
        p.max = mx;
        p.cnt = cnt;
        p.pad = 0ull;
        parts[slot] = p;
    }
}
