// kf_reduce.hpp -- wavefront/block reductions and the last-block finalize of a
// deform iteration (FP64 max |du|, FP64 energy sum in a fixed order, active count).
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

// Publish this block's partials; the block that arrives last folds every block's
// energy partial in index order, finalises the iteration into st and re-arms the
// counters for the next launch.  Must be called by all threads of the block.
template <int BLOCK>
__device__ inline void kf_block_finalize(DeformState *st, double *partials, unsigned int nblocks,
                                         unsigned int block_id, double sum, double mx,
                                         unsigned long long cnt)
{
    __shared__ bool s_last;
    if (threadIdx.x == 0) {
        partials[block_id] = sum;
        if (mx > 0.0)
            atomicMax(&st->max_bits, (unsigned long long)__double_as_longlong(mx));
        if (cnt)
            atomicAdd(&st->active, cnt);
        __threadfence();
        unsigned int t = atomicAdd(&st->ticket, 1u);
        s_last = (t == nblocks - 1);
    }
    __syncthreads();
    if (!s_last)
        return;
    __threadfence();
    double acc = 0.0;
    for (unsigned int b = threadIdx.x; b < nblocks; b += BLOCK)
        acc += __builtin_nontemporal_load(&partials[b]);
    double dummy_m = 0.0;
    unsigned long long dummy_c = 0;
    kf_block_reduce3<BLOCK>(acc, dummy_m, dummy_c);
    if (threadIdx.x == 0) {
        unsigned long long mb = atomicAdd(&st->max_bits, 0ull);
        unsigned long long act = atomicAdd(&st->active, 0ull);
        double m = __longlong_as_double((long long)mb);
        st->last_max = m;
        st->energy = acc;
        st->last_active = act;
        st->iters += 1;
        long long dv = atomicAdd((unsigned long long *)&st->diverged, 0ull);
        if (dv != LLONG_MAX || (st->stop > 0.0 && m < st->stop))
            st->done = 1;
        st->max_bits = 0ull;
        st->active = 0ull;
        st->ticket = 0u;
        __threadfence();
    }
}
