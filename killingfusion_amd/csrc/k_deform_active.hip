// k_deform_active.hip -- the adjoint Jacobi step over the voxels that can move.
//
// A step moves voxel j only when it is active, -clip <= psi_j <= Max - eps with psi_j the live
// SDF at j + u_j (KillingFusion.cpp:272-274, 317-319: the update lands only on voxels that
// pass the test).  With u_j = 0 the trilinear read has zero weights and returns phi_live[j]
// exactly, so a voxel with u_j = 0 and phi_live[j] outside the band is inactive, keeps
// u_j = 0 and, by induction, never moves.  The voxels that can move during one deform call
// are therefore fixed when it starts:
//     C = { j : some component of u_j has a nonzero bit pattern, or -clip <= phi_live[j] <= Max - eps }.
// (-0 counts as nonzero here, so every voxel outside C holds +0 in all three components.)
//
// Per call:
//  * k_active_mark streams u and phi_live once: one bit per voxel (64-voxel words) and a count
//    per span of 4096 voxels; k_active_scan turns the counts into span offsets and the total.
//  * k_active_compact lists C in ascending voxel order (one slot per listed voxel);
//    k_active_neighbours copies each listed u_j into a compact array (one float4 per slot) and
//    finds the slots of the 12 neighbours the Killing term reads (-1: outside the grid or not
//    listed, i.e. +0).
//  * k_adj_list is one iteration: it reads a listed voxel's own u and its neighbours' from the
//    compact array (about a megabyte that stays in L2), gathers the live SDF around j + u_j,
//    runs the same per-voxel arithmetic as the dense step (adj_voxel, adj_update) and writes
//    the new u_j into the other compact set.
//  * k_active_energy evaluates the energy of the final field over the listed voxels and the
//    unlisted ones next to them (every other voxel's term is zero), with adj_voxel's
//    per-voxel terms; k_active_scatter writes the final compact values back into the field.
// Every iteration therefore computes the dense step's values bit for bit; the final energy
// has the same per-voxel terms, summed in FP64 in another order.
#include "kf_adjoint.hpp"

#include <cstdio>

#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int LB = 64;                 // threads per block of the list step (one wavefront)
constexpr int SPAN = BLOCK * 16;       // voxels per block of the marking and listing passes
constexpr int WPB = SPAN / 64;         // 64-voxel mask words per such block (one per thread)
constexpr int NNB = 12;                // neighbours of the Killing term: 6 faces, 6 edges
constexpr int SCAN_BLOCK = 1024;
// The single-volume list kernels run over a number of slots that is a multiple of this (so of
// every block size they use); slots past the list's length do nothing.
constexpr unsigned int SLOT_QUANTUM = 1024;

// Neighbour k of voxel (x, y, z): faces +x -x +y -y +z -z, then the edge neighbours of Nbhd::dgl
// in the order dgl[0][1], dgl[0][2], dgl[1][0], dgl[1][2], dgl[2][0], dgl[2][1].
__constant__ int c_nb_off[NNB][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                                     {1, -1, 0}, {1, 0, -1}, {-1, 1, 0}, {0, 1, -1}, {-1, 0, 1}, {0, -1, 1}};

// READ_U false: every u is +0 (the caller has just cleared the field).  *nonfinite is set when
// some live SDF value is not finite.
// Spreads the 16 bits of x to bits 0, 4, .., 60.
__device__ inline unsigned long long spread4(unsigned long long x)
{
    x &= 0xFFFFull;
    x = (x | (x << 24)) & 0x000000FF000000FFull;
    x = (x | (x << 12)) & 0x000F000F000F000Full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
    return x;
}

template <class S, bool READ_U>
__device__ inline void mark_span(const PhiRead<S> &L, const float *__restrict__ u, float lo, float hi,
                                 unsigned long long *__restrict__ mask, unsigned int *__restrict__ span_count,
                                 unsigned int *nonfinite, unsigned int sb)
{
    const unsigned int n32 = nvox32(L.g);
    const unsigned int base = sb * (unsigned int)SPAN;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int cnt = 0;
    if (n32 % 4u == 0u) {
        // Four voxels per lane (one float4 of each u plane): a wavefront covers 256 voxels, four
        // mask words.  Ballot v holds voxel 4 lane + v of each lane; word j takes lanes
        // 16 j .. 16 j + 15, bit 4 m + v from bit 16 j + m of ballot v.  Lanes 0 .. 3 keep their
        // words in registers and store them (and the non-finite flag) after the loop: one memory
        // counter tracks loads and stores here, so a store inside the loop would make every later
        // round's loads wait for it.
        constexpr int ROUNDS = SPAN / (BLOCK * 4);
        unsigned long long words[ROUNDS];
        bool any_bad = false;
        // Every round's loads go out before any is used.  A round past the grid's end reads the
        // group at 0 instead (in bounds: N is a nonzero multiple of 4 here) and is ignored.
        float live[ROUNDS][4];
        float4 q[ROUNDS][3];
#pragma unroll
        for (int k = 0; k < ROUNDS; k++) {
            const unsigned int i = base + (unsigned int)(k * BLOCK * 4 + w * 256) + 4u * (unsigned int)lane;
            const unsigned int ic = i < n32 ? i : 0u;
            S::load4(L.p, (int64_t)ic, L.scale, live[k]);
#pragma unroll
            for (int a = 0; a < 3; a++)
                q[k][a] = READ_U ? kf_load_stream4(u + (size_t)a * n32 + ic) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < ROUNDS; k++) {
            const unsigned int w0 = base + (unsigned int)(k * BLOCK * 4 + w * 256);
            const unsigned int i = w0 + 4u * (unsigned int)lane;
            bool cand[4] = {false, false, false, false}, bad = false;
            if (i < n32) {
                unsigned int ub[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    ub[0] |= __float_as_uint(q[k][a].x);
                    ub[1] |= __float_as_uint(q[k][a].y);
                    ub[2] |= __float_as_uint(q[k][a].z);
                    ub[3] |= __float_as_uint(q[k][a].w);
                }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    bad = bad || !isfinite(live[k][v]);
                    cand[v] = ub[v] != 0u || !(live[k][v] > hi || live[k][v] < lo);
                }
            }
            const bool wave_bad = __ballot(bad) != 0ull;  // every lane runs the ballot
            any_bad = any_bad || wave_bad;
            unsigned long long bal[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                bal[v] = __ballot(cand[v]);
                cnt += (unsigned int)__popcll(bal[v]);
            }
            words[k] = 0ull;
#pragma unroll
            for (int v = 0; v < 4; v++)
                words[k] |= spread4(bal[v] >> (16 * (lane & 3))) << v;
        }
#pragma unroll
        for (int k = 0; k < ROUNDS; k++) {
            const unsigned int w0 = base + (unsigned int)(k * BLOCK * 4 + w * 256);
            if (lane < 4 && w0 + 64u * (unsigned int)lane < n32)
                mask[(w0 >> 6) + (unsigned int)lane] = words[k];
        }
        if (any_bad && lane == 0)
            atomicOr(nonfinite, 1u);
    } else {
        // One voxel per lane, 16 rounds taken 4 at a time: a group's loads go out before any is
        // used (a voxel past the grid's end reads voxel 0 instead and is ignored), and the mask
        // words and the non-finite flag are stored after the loop, as above.
        constexpr int ROUNDS = SPAN / BLOCK, GROUP = 4;
        unsigned long long words[ROUNDS];
        bool any_bad = false;
#pragma unroll
        for (int k0 = 0; k0 < ROUNDS; k0 += GROUP) {
            float live[GROUP];
            unsigned int ub[GROUP];
#pragma unroll
            for (int k = 0; k < GROUP; k++) {
                const unsigned int i = base + (unsigned int)((k0 + k) * BLOCK + w * 64) + (unsigned int)lane;
                const unsigned int ic = i < n32 ? i : 0u;
                ub[k] = READ_U ? __float_as_uint(kf_load_stream(u + ic)) |
                                     __float_as_uint(kf_load_stream(u + (size_t)n32 + ic)) |
                                     __float_as_uint(kf_load_stream(u + 2 * (size_t)n32 + ic))
                               : 0u;
                live[k] = S::load(L.p, (int64_t)ic, L.scale);
            }
#pragma unroll
            for (int k = 0; k < GROUP; k++) {
                const unsigned int i = base + (unsigned int)((k0 + k) * BLOCK + w * 64) + (unsigned int)lane;
                bool cand = false, bad = false;
                if (i < n32) {
                    bad = !isfinite(live[k]);
                    cand = ub[k] != 0u || !(live[k] > hi || live[k] < lo);
                }
                const bool wave_bad = __ballot(bad) != 0ull;  // every lane runs the ballot
                any_bad = any_bad || wave_bad;
                words[k0 + k] = __ballot(cand);
                cnt += (unsigned int)__popcll(words[k0 + k]);
            }
        }
#pragma unroll
        for (int k = 0; k < ROUNDS; k++) {
            const unsigned int w0 = base + (unsigned int)(k * BLOCK + w * 64);
            if (lane == 0 && w0 < n32)
                mask[w0 >> 6] = words[k];
        }
        if (any_bad && lane == 0)
            atomicOr(nonfinite, 1u);
    }
    __shared__ unsigned int s_cnt[BLOCK / 64];
    if (lane == 0)
        s_cnt[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int k = 0; k < BLOCK / 64; k++)
            t += s_cnt[k];
        span_count[sb] = t;
    }
}

template <class S, bool READ_U>
__global__ __launch_bounds__(BLOCK) void k_active_mark(PhiRead<S> L, const float *__restrict__ u, float lo, float hi,
                                                       unsigned long long *__restrict__ mask,
                                                       unsigned int *__restrict__ span_count, unsigned int *nonfinite)
{
    mark_span<S, READ_U>(L, u, lo, hi, mask, span_count, nonfinite, blockIdx.x);
}

// Exclusive scan of the span counts (one block); the total goes to *total.  The block walks
// the counts in tiles of 8 SCAN_BLOCK, a thread eight consecutive counts of each (loaded
// together), carrying the running total from tile to tile (712^3: 88 k spans, 162 -> 84 us;
// 32 counts per thread measured the same).
__device__ inline void scan_spans(const unsigned int *__restrict__ span_count, unsigned int nspan,
                                  unsigned int *__restrict__ span_off, unsigned int *total)
{
    constexpr unsigned int PT = 8, TILE = PT * SCAN_BLOCK;
    __shared__ unsigned int s_w[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int carry = 0;  // the same in every thread
    for (unsigned int t0 = 0; t0 < nspan; t0 += TILE) {
        const unsigned int b0 = t0 + threadIdx.x * PT;
        unsigned int q[PT], sum = 0;
#pragma unroll
        for (unsigned int j = 0; j < PT; j++) {
            q[j] = b0 + j < nspan ? span_count[b0 + j] : 0u;
            sum += q[j];
        }
        unsigned int v = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned int t = __shfl_up(v, o, 64);
            if (lane >= o)
                v += t;
        }
        if (lane == 63)
            s_w[w] = v;
        __syncthreads();
        unsigned int before = 0, tile_total = 0;
        for (int k = 0; k < SCAN_BLOCK / 64; k++) {
            const unsigned int t = s_w[k];
            before += k < w ? t : 0u;
            tile_total += t;
        }
        unsigned int run = carry + before + v - sum;
#pragma unroll
        for (unsigned int j = 0; j < PT; j++)
            if (b0 + j < nspan) {
                span_off[b0 + j] = run;
                run += q[j];
            }
        carry += tile_total;
        __syncthreads();  // s_w is rewritten by the next tile
    }
    if (threadIdx.x == SCAN_BLOCK - 1)
        *total = carry;
}

// The single-volume scan also decides, on the device, whether the list kernels run: they were
// launched for `cap` slots, so a longer list (or, with need_finite, a non-finite live value
// seen by the marking pass) sets list_bad and list_n = 0, and every later list kernel of the
// call does nothing.  The host reads list_bad with the call's statistics.
__global__ __launch_bounds__(SCAN_BLOCK) void k_active_scan(const unsigned int *__restrict__ span_count,
                                                            unsigned int nspan, unsigned int *__restrict__ span_off,
                                                            unsigned int *total)
{
    scan_spans(span_count, nspan, span_off, total);
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_active_scan_sized(const unsigned int *__restrict__ span_count,
                                                                  unsigned int nspan,
                                                                  unsigned int *__restrict__ span_off, DeformState *st,
                                                                  unsigned int cap, int need_finite)
{
    scan_spans(span_count, nspan, span_off, &st->listed);
    if (threadIdx.x == SCAN_BLOCK - 1) {
        const unsigned int total = st->listed;  // this thread's own store
        const bool bad = total > cap || (need_finite && st->nonfinite != 0u);
        st->list_n = bad ? 0u : total;
        st->list_bad = bad ? 1u : 0u;
    }
}

// The voxels a warp (SDF::update) can change: u_j != 0 (any component; -0 counts as 0, it
// moves nothing), plus voxels whose phi is -0 or not finite.  Elsewhere the warp reads its own
// cell with zero weights and returns phi and w unchanged, provided every phi is finite (a
// non-finite corner times a zero weight is NaN): a non-finite phi sets *flag.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_moved_mark(const typename S::T *__restrict__ phi, float scale,
                                                      const float *__restrict__ u, unsigned int n32,
                                                      unsigned long long *__restrict__ mask,
                                                      unsigned int *__restrict__ span_count, unsigned int *flag)
{
    const unsigned int base = blockIdx.x * (unsigned int)SPAN;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int cnt = 0;
#pragma unroll 4
    for (int k = 0; k < SPAN / BLOCK; k++) {
        const unsigned int w0 = base + (unsigned int)(k * BLOCK + w * 64);
        const unsigned int i = w0 + (unsigned int)lane;
        bool cand = false, bad = false;
        if (i < n32) {
            const float u0 = kf_load_stream(u + i), u1 = kf_load_stream(u + (size_t)n32 + i);
            const float u2 = kf_load_stream(u + 2 * (size_t)n32 + i);
            const float f = S::load(phi, (int64_t)i, scale);
            bad = !isfinite(f);
            cand = ((__float_as_uint(u0) | __float_as_uint(u1) | __float_as_uint(u2)) & 0x7fffffffu) != 0u ||
                   __float_as_uint(f) == 0x80000000u || bad;
        }
        if (__ballot(bad) != 0ull && lane == 0)
            atomicOr(flag, 1u);
        const unsigned long long bal = __ballot(cand);
        if (lane == 0 && w0 < n32)
            mask[w0 >> 6] = bal;
        cnt += (unsigned int)__popcll(bal);
    }
    __shared__ unsigned int s_cnt[BLOCK / 64];
    if (lane == 0)
        s_cnt[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int k = 0; k < BLOCK / 64; k++)
            t += s_cnt[k];
        span_count[blockIdx.x] = t;
    }
}

__global__ void k_clear_words(unsigned int *p, int n)
{
    if ((int)threadIdx.x < n)
        p[threadIdx.x] = 0u;
}

// Thread k of span block sb lists the voxels of mask word sb WPB + k, in voxel order.
__device__ inline void compact_span(const unsigned long long *__restrict__ mask, unsigned int nwords,
                                    const unsigned int *__restrict__ span_count,
                                    const unsigned int *__restrict__ span_off, unsigned int base,
                                    unsigned int *__restrict__ list, unsigned int *__restrict__ pos_of, unsigned int sb)
{
    if (span_count[sb] == 0u)
        return;
    const unsigned int wd = sb * (unsigned int)WPB + threadIdx.x;
    unsigned long long m = wd < nwords ? mask[wd] : 0ull;
    const unsigned int c = (unsigned int)__popcll(m);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int v = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int t = __shfl_up(v, o, 64);
        if (lane >= o)
            v += t;
    }
    __shared__ unsigned int s_w[WPB / 64];
    if (lane == 63)
        s_w[w] = v;
    __syncthreads();
    unsigned int wbase = 0;
    for (int k = 0; k < w; k++)
        wbase += s_w[k];
    unsigned int pos = base + span_off[sb] + wbase + v - c;
    while (m) {
        const unsigned int i = wd * 64u + (unsigned int)__builtin_ctzll(m);
        m &= m - 1ull;
        list[pos] = i;
        pos_of[i] = pos;
        pos++;
    }
}

__global__ __launch_bounds__(WPB) void k_active_compact(const unsigned long long *__restrict__ mask, unsigned int nwords,
                                                        const unsigned int *__restrict__ span_count,
                                                        const unsigned int *__restrict__ span_off, unsigned int base,
                                                        unsigned int *__restrict__ list, unsigned int *__restrict__ pos_of)
{
    compact_span(mask, nwords, span_count, span_off, base, list, pos_of, blockIdx.x);
}

__global__ __launch_bounds__(WPB) void k_active_compact_sized(const unsigned long long *__restrict__ mask,
                                                              unsigned int nwords,
                                                              const unsigned int *__restrict__ span_count,
                                                              const unsigned int *__restrict__ span_off,
                                                              const DeformState *st, unsigned int *__restrict__ list,
                                                              unsigned int *__restrict__ pos_of)
{
    if (st->list_bad != 0u)
        return;  // the list does not fit the slots it was given
    compact_span(mask, nwords, span_count, span_off, 0u, list, pos_of, blockIdx.x);
}

// Slot t (base .. base + count - 1): u of the listed voxel (compact set 0) and the list slots
// of its 12 neighbours.
// (Thread of local block lb.)
__device__ inline void list_neighbours(const GridDims &g, const unsigned long long *__restrict__ mask,
                                       const unsigned int *__restrict__ pos_of, const unsigned int *__restrict__ list,
                                       unsigned int base, unsigned int count, const float *__restrict__ u,
                                       float4 *__restrict__ uc, int4 *__restrict__ nbr, unsigned int lb)
{
    if (lb * (unsigned int)BLOCK + threadIdx.x >= count)
        return;
    const unsigned int t = base + lb * (unsigned int)BLOCK + threadIdx.x;
    const unsigned int i = list[t];
    const unsigned int n32 = nvox32(g);
    uc[t] = make_float4(u[i], u[(size_t)n32 + i], u[2 * (size_t)n32 + i], 0.0f);
    int x, y, z;
    g.xyz32(i, x, y, z);
    int slot[NNB];
#pragma unroll
    for (int k = 0; k < NNB; k++) {
        const int xn = x + c_nb_off[k][0], yn = y + c_nb_off[k][1], zn = z + c_nb_off[k][2];
        slot[k] = -1;
        if (g.in(xn, yn, zn)) {
            const unsigned int j = g.idx32(xn, yn, zn);
            if ((mask[j >> 6] >> (j & 63u)) & 1ull)
                slot[k] = (int)pos_of[j];
        }
    }
#pragma unroll
    for (int q = 0; q < NNB / 4; q++)
        nbr[3u * t + (unsigned int)q] = make_int4(slot[4 * q], slot[4 * q + 1], slot[4 * q + 2], slot[4 * q + 3]);
}

__global__ __launch_bounds__(BLOCK) void k_active_neighbours(GridDims g, const unsigned long long *__restrict__ mask,
                                                             const unsigned int *__restrict__ pos_of,
                                                             const unsigned int *__restrict__ list,
                                                             const DeformState *st, const float *__restrict__ u,
                                                             float4 *__restrict__ uc, int4 *__restrict__ nbr)
{
    list_neighbours(g, mask, pos_of, list, 0u, st->list_n, u, uc, nbr, blockIdx.x);
}

// Listing of a batch of volumes, four launches for all of them.  Volume v's spans are blocks
// sb0 .. sb0 + nspan - 1 of the marking and compaction launches and its listed voxels blocks
// nb0 .. of the neighbours launch (span_vol / nb_vol map a block to its volume).  Each volume
// has its own mask words, span counts and offsets, and pos_of entries.
template <class S>
struct ListVol {
    PhiRead<S> L;
    const float *u;
    unsigned long long *mask;
    unsigned int *span_count;  // nspan counts, then nspan offsets
    unsigned int *pos_of;
    unsigned int *nonfinite, *listed;
    unsigned int nspan, sb0, base, cnt, nb0;
};

template <class S>
__global__ __launch_bounds__(BLOCK) void k_active_mark_batch(const ListVol<S> *__restrict__ vols,
                                                             const unsigned int *__restrict__ span_vol, float lo,
                                                             float hi)
{
    const ListVol<S> &V = vols[span_vol[blockIdx.x]];
    mark_span<S, true>(V.L, V.u, lo, hi, V.mask, V.span_count, V.nonfinite, blockIdx.x - V.sb0);
}

template <class S>
__global__ __launch_bounds__(SCAN_BLOCK) void k_active_scan_batch(const ListVol<S> *__restrict__ vols)
{
    const ListVol<S> &V = vols[blockIdx.x];
    scan_spans(V.span_count, V.nspan, V.span_count + V.nspan, V.listed);
}

template <class S>
__global__ __launch_bounds__(WPB) void k_active_compact_batch(const ListVol<S> *__restrict__ vols,
                                                              const unsigned int *__restrict__ span_vol,
                                                              unsigned int *__restrict__ list)
{
    const ListVol<S> &V = vols[span_vol[blockIdx.x]];
    compact_span(V.mask, (unsigned int)((V.L.g.n + 63) / 64), V.span_count, V.span_count + V.nspan, V.base, list,
                 V.pos_of, blockIdx.x - V.sb0);
}

template <class S>
__global__ __launch_bounds__(BLOCK) void k_active_neighbours_batch(const ListVol<S> *__restrict__ vols,
                                                                   const unsigned int *__restrict__ nb_vol,
                                                                   const unsigned int *__restrict__ list,
                                                                   float4 *__restrict__ uc, int4 *__restrict__ nbr)
{
    const ListVol<S> &V = vols[nb_vol[blockIdx.x]];
    list_neighbours(V.L.g, V.mask, V.pos_of, list, V.base, V.cnt, V.u, uc, nbr, blockIdx.x - V.nb0);
}

// Fold of step launch k - 1 by block 0 of launch k: kf_fold_launch's result, with every load
// (the gate flag, the state fields the fold reads and the partials) issued before any of them
// is needed, so the fold costs one round trip.  Partials are read even when the gate is
// closed; their values are then ignored.  A launch that published no partials (prev_nb = 0:
// no stop rule, not the last launch) still counts as an iteration and reports its divergence;
// its max |du| and active count read as 0 until the last launch's are folded.
template <int NT>
__device__ inline void list_fold(DeformState *st, const BlockPartial *prev, unsigned int prev_nb, bool has_prev,
                                 int parity)
{
    const unsigned int go = st->run[parity];
    const int pp = (parity ^ 1) & 1;
    unsigned int done = 0u, iters = 0u;
    long long dv = LLONG_MAX;
    double stop = 0.0;
    if (threadIdx.x == 0) {
        done = st->done;
        iters = st->iters;
        dv = st->div_slot[pp];
        stop = st->stop;
    }
    double s = 0.0, m = 0.0;
    unsigned long long c = 0;
    for (unsigned int b0 = threadIdx.x; b0 < prev_nb; b0 += 8 * NT) {
        BlockPartial p[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const unsigned int b = b0 + (unsigned int)j * NT;
            p[j] = b < prev_nb ? prev[b] : BlockPartial{0.0, 0.0, 0ull, 0ull};
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            s += p[j].sum;
            m = fmax(m, p[j].max);
            c += p[j].cnt;
        }
    }
    kf_block_reduce3<NT>(s, m, c);
    if (threadIdx.x == 0) {
        if (go && has_prev && !done) {
            st->div_slot[pp] = LLONG_MAX;
            st->last_max = m;
            st->energy = s;
            st->last_active = c;
            st->iters = iters + 1u;
            if (dv != LLONG_MAX)
                st->diverged = dv;
            if (dv != LLONG_MAX || (stop > 0.0 && m < stop)) {
                st->done = 1u;
                done = 1u;
            }
        }
        st->run[parity ^ 1] = (go && !done) ? 1u : 0u;
    }
}

__device__ inline float comp(const float4 &v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }

// The neighbourhood of listed slot t from compact set `uc`: faces (three components each)
// and the edge neighbours of Nbhd::dgl (one component each); slot -1 reads +0.
__device__ inline void list_nbhd(const float4 *__restrict__ uc, const int ns[NNB], Nbhd &nb)
{
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 f[6];
#pragma unroll
    for (int k = 0; k < 6; k++)
        f[k] = ns[k] >= 0 ? uc[ns[k]] : zero;
    const float *ucf = reinterpret_cast<const float *>(uc);
    constexpr int ecomp[6] = {1, 2, 0, 2, 0, 1};
    float e[6];
#pragma unroll
    for (int k = 0; k < 6; k++)
        e[k] = ns[6 + k] >= 0 ? ucf[4u * (unsigned int)ns[6 + k] + (unsigned int)ecomp[k]] : 0.0f;
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            nb.up[b][a] = comp(f[2 * b], a);
            nb.um[b][a] = comp(f[2 * b + 1], a);
        }
    nb.dgl[0][1] = e[0];
    nb.dgl[0][2] = e[1];
    nb.dgl[1][0] = e[2];
    nb.dgl[1][2] = e[3];
    nb.dgl[2][0] = e[4];
    nb.dgl[2][1] = e[5];
}

__device__ inline void load_slots(const int4 *__restrict__ nbr, unsigned int t, int ns[NNB])
{
#pragma unroll
    for (int q = 0; q < NNB / 4; q++) {
        const int4 v = nbr[3u * t + (unsigned int)q];
        ns[4 * q] = v.x;
        ns[4 * q + 1] = v.y;
        ns[4 * q + 2] = v.z;
        ns[4 * q + 3] = v.w;
    }
}

__device__ inline void voxel_xyz(const GridDims &g, unsigned int i, int &x, int &y, int &z)
{
    g.xyz32(i, x, y, z);
}

// Round 2 and the update of listed slot t of voxel i, given its own u and neighbour slots:
// the live cell around i + u_i, phi_can[i], the neighbours' u, then adj_voxel / adj_update.
template <class S>
__device__ inline void list_step(const PhiRead<S> &L, const PhiRead<S> &Cn, const float4 *__restrict__ uin,
                                 float4 *__restrict__ uout, const AdjConsts &c, long long *div, unsigned int t,
                                 unsigned int i, const float4 &own, const int ns[NNB], double &mx,
                                 unsigned long long &act)
{
    int x, y, z, bx, by, bz;
    voxel_xyz(L.g, i, x, y, z);
    Nbhd nb;
    nb.uc[0] = own.x;
    nb.uc[1] = own.y;
    nb.uc[2] = own.z;
    float fx, fy, fz, v[8];
    kf_split(x, nb.uc[0], bx, fx);
    kf_split(y, nb.uc[1], by, fy);
    kf_split(z, nb.uc[2], bz, fz);
    cell32(L, bx, by, bz, v);
    const float can = c.use_data ? S::load(Cn.p, (int64_t)i, Cn.scale) : 0.f;
    list_nbhd(uin, ns, nb);
    VoxelOut o;
    adj_voxel<false>(c, v, fx, fy, fz, can, nb, x, y, z, o);
    if (o.active)
        act = 1;
    float n[3];
    adj_update(c, o, nb.uc, (int64_t)i, div, mx, n);
    uout[t] = make_float4(n[0], n[1], n[2], 0.0f);
}

// Energy of field `uc` (see the file header of k_deform_adj.hip for E) owned by slot t: the
// term of its listed voxel j and of each unlisted voxel i = j - e_b (in the grid) for which j
// is the first listed voxel among i + e_0, i + e_1, i + e_2 -- so every unlisted voxel next to
// a listed one is counted once.  An unlisted voxel holds u = +0 and phi_live[i] outside the
// band (it is inactive), so only its Killing term is nonzero, and that reads u at i + e_b only.
template <class S>
__device__ inline double list_energy(const PhiRead<S> &L, const PhiRead<S> &Cn, const float4 *__restrict__ uc,
                                     const AdjConsts &c, unsigned int t, unsigned int i, const float4 &own,
                                     const int ns[NNB])
{
    int x, y, z, bx, by, bz;
    voxel_xyz(L.g, i, x, y, z);
    Nbhd nb;
    nb.uc[0] = own.x;
    nb.uc[1] = own.y;
    nb.uc[2] = own.z;
    float fx, fy, fz, v[8];
    kf_split(x, nb.uc[0], bx, fx);
    kf_split(y, nb.uc[1], by, fy);
    kf_split(z, nb.uc[2], bz, fz);
    cell32(L, bx, by, bz, v);
    const float can = c.use_data ? S::load(Cn.p, (int64_t)i, Cn.scale) : 0.f;
    list_nbhd(uc, ns, nb);
    VoxelOut o;
    adj_voxel<true>(c, v, fx, fy, fz, can, nb, x, y, z, o);
    double esum = (double)o.e;
    // Slots: faces -x -y -z = 1, 3, 5; the edge slots i - e_b + e_b' are
    // (b, b') = (1, 0): 6, (2, 0): 7, (2, 1): 9, (0, 1): 8, (0, 2): 10, (1, 2): 11.
    const int jj[3] = {x, y, z};
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int b = 0; b < 3; b++) {
        if (jj[b] == 0 || ns[2 * b + 1] >= 0)
            continue;  // i outside the grid, or listed
        const bool earlier = b == 1 ? ns[6] >= 0 : b == 2 ? (ns[7] >= 0 || ns[9] >= 0) : false;
        if (earlier)
            continue;  // counted by the listed voxel i + e_b', b' < b
        // u at i + e_0, i + e_1, i + e_2
        int sl[3];
        if (b == 0) {
            sl[0] = (int)t; sl[1] = ns[8]; sl[2] = ns[10];
        } else if (b == 1) {
            sl[0] = ns[6]; sl[1] = (int)t; sl[2] = ns[11];
        } else {
            sl[0] = ns[7]; sl[1] = ns[9]; sl[2] = (int)t;
        }
        Nbhd ni;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float4 q = sl[d] >= 0 ? uc[sl[d]] : zero;
            ni.uc[d] = 0.0f;
            ni.up[d][0] = q.x;
            ni.up[d][1] = q.y;
            ni.up[d][2] = q.z;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                ni.um[d][a] = 0.0f;
                ni.dgl[d][a] = 0.0f;
            }
        }
        const int xi = x - (b == 0), yi = y - (b == 1), zi = z - (b == 2);
        const float li = S::load(L.p, (int64_t)L.g.idx32(xi, yi, zi), L.scale);
        float vi[8];
#pragma unroll
        for (int k = 0; k < 8; k++)
            vi[k] = li;
        VoxelOut oi;
        adj_voxel<true>(c, vi, 0.0f, 0.0f, 0.0f, 0.0f, ni, xi, yi, zi, oi);
        esum += (double)oi.e;
    }
    return esum;
}

// Work block w -> block of list slots.  The first nmap blocks (the list's expected length) are
// dealt so that consecutive runs of the list (neighbouring voxels) go to one XCD and every
// XCD gets an equal share; blocks past them (slots the list may not reach) keep their index.
__device__ inline unsigned int list_tile(unsigned int w, unsigned int nmap)
{
    return w < nmap ? kf_xcd_tile(w, nmap) : w;
}

// One Jacobi iteration over the list.  Block 0 only folds the previous launch and arms the
// next one (step_gate's first-block duty), beside the stencils rather than in front of one.
// A work block loads its list entries, own u and neighbour slots together with the gate flag
// (one round trip), then the live cell and the neighbours' u (a second one).
template <class S>
__global__ __launch_bounds__(LB) void k_adj_list(PhiRead<S> L, PhiRead<S> Cn, const float4 *__restrict__ uin,
                                                 float4 *__restrict__ uout, AdjConsts c, DeformState *st,
                                                 BlockPartial *partials, int parity, const BlockPartial *prev,
                                                 unsigned int prev_nb, int has_prev, int publish,
                                                 const unsigned int *__restrict__ list,
                                                 const int4 *__restrict__ nbr, unsigned int nmap)
{
    if (blockIdx.x == 0) {
        list_fold<LB>(st, prev, prev_nb, has_prev != 0, parity);
        return;
    }
    const unsigned int w = blockIdx.x - 1u;
    if (w >= nmap && w * (unsigned int)LB >= st->list_n) {
        // slots past the expected length and past the list's end: only the (empty) partial
        if (publish && st->run[parity] != 0u)
            kf_block_publish<LB>(partials, w, 0.0, 0.0, 0ull);
        return;
    }
    const unsigned int t = list_tile(w, nmap) * (unsigned int)LB + threadIdx.x;
    // Every slot of the grid lies within the list's capacity, so the entries are loaded with
    // the list length (on the device) in one round trip and used only below it.
    const unsigned int i = list[t];
    const float4 own = uin[t];
    int ns[NNB];
    load_slots(nbr, t, ns);
    const unsigned int count = st->list_n;
    if (st->run[parity] == 0u)
        return;
    const bool mine = t < count;
    double mx = 0.0;
    unsigned long long act = 0;
    if (mine)
        list_step(L, Cn, uin, uout, c, &st->div_slot[parity], t, i, own, ns, mx, act);
    if (publish)
        kf_block_publish<LB>(partials, w, 0.0, mx, act);
}

// Energy of the final compact set (st->iters & 1) into one partial per block.
template <class S>
__global__ __launch_bounds__(LB) void k_active_energy(PhiRead<S> L, PhiRead<S> Cn, const float4 *__restrict__ u0,
                                                      unsigned int set, const DeformState *st, AdjConsts c,
                                                      BlockPartial *partials, const unsigned int *__restrict__ list,
                                                      const int4 *__restrict__ nbr, unsigned int nmap)
{
    const float4 *uc = u0 + (size_t)(st->iters & 1u) * set;
    const unsigned int t = list_tile(blockIdx.x, nmap) * (unsigned int)LB + threadIdx.x;
    double esum = 0.0;
    if (t < st->list_n) {
        int ns[NNB];
        load_slots(nbr, t, ns);
        esum = list_energy(L, Cn, uc, c, t, list[t], uc[t], ns);
    }
    kf_block_publish<LB>(partials, blockIdx.x, esum, 0.0, 0ull);
}

// The energy partials in slot order -> st->list_energy.
__global__ __launch_bounds__(1024) void k_list_energy_fold(DeformState *st, const BlockPartial *__restrict__ parts,
                                                           unsigned int nb)
{
    double s = 0.0, m = 0.0;
    unsigned long long c = 0;
    for (unsigned int b = threadIdx.x; b < nb; b += 1024)
        s += parts[b].sum;
    kf_block_reduce3<1024>(s, m, c);
    if (threadIdx.x == 0)
        st->list_energy = s;
}

// ---------------------------------------------------------------------------
// Batched active-set sweep: the lists of several volumes in one slot space.  Volume v owns
// work blocks blk0 .. blk0 + nblk - 1 and their slots (LB per block; slots past the volume's
// count hold the index NONE); each volume has its own iteration state, gate and partials.
// ---------------------------------------------------------------------------
constexpr unsigned int NONE = 0xffffffffu;

template <class S>
struct ActiveVol {
    PhiRead<S> L, Cn;
    float *u;            // the field's current buffer (SoA [3][n])
    unsigned int blk0;   // first work block
    unsigned int nblk;   // work blocks (at least one)
};

// Launch k: blocks 0 .. nvol - 1 fold launch k - 1 of their volume (has_prev; its partials
// when it published them) and arm launch k + 1; the other blocks are work blocks, which
// publish partials when `publish` is set.
template <class S>
__global__ __launch_bounds__(LB) void k_adj_list_batch(const ActiveVol<S> *__restrict__ vols, int nvol,
                                                       const unsigned int *__restrict__ blk_vol,
                                                       const float4 *__restrict__ uin, float4 *__restrict__ uout,
                                                       AdjConsts c, DeformState *states, BlockPartial *partials,
                                                       int parity, const BlockPartial *prev, int has_prev,
                                                       int prev_published, int publish,
                                                       const unsigned int *__restrict__ list,
                                                       const int4 *__restrict__ nbr)
{
    if ((int)blockIdx.x < nvol) {
        const unsigned int b0 = vols[blockIdx.x].blk0, nb = vols[blockIdx.x].nblk;
        list_fold<LB>(states + blockIdx.x, prev + b0, prev_published ? nb : 0u, has_prev != 0, parity);
        return;
    }
    const unsigned int w = kf_xcd_tile(blockIdx.x - (unsigned int)nvol, gridDim.x - (unsigned int)nvol);
    const unsigned int t = w * (unsigned int)LB + threadIdx.x;
    const unsigned int i = list[t];
    const float4 own = uin[t];
    int ns[NNB];
    load_slots(nbr, t, ns);
    const unsigned int v = blk_vol[w];
    DeformState *st = states + v;
    if (st->run[parity] == 0u)
        return;
    double mx = 0.0;
    unsigned long long act = 0;
    if (i != NONE)
        list_step(vols[v].L, vols[v].Cn, uin, uout, c, &st->div_slot[parity], t, i, own, ns, mx, act);
    if (publish)
        kf_block_publish<LB>(partials, w, 0.0, mx, act);
}

// Energy of each volume's final compact set (iters & 1 of its state) into partials, one per
// work block.
template <class S>
__global__ __launch_bounds__(LB) void k_active_energy_batch(const ActiveVol<S> *__restrict__ vols,
                                                            const unsigned int *__restrict__ blk_vol,
                                                            const float4 *__restrict__ u0, unsigned int set,
                                                            AdjConsts c, const DeformState *states,
                                                            BlockPartial *partials, const unsigned int *__restrict__ list,
                                                            const int4 *__restrict__ nbr)
{
    const unsigned int w = kf_xcd_tile(blockIdx.x, gridDim.x);
    const unsigned int t = w * (unsigned int)LB + threadIdx.x;
    const unsigned int v = blk_vol[w];
    const float4 *uc = u0 + (size_t)(states[v].iters & 1u) * set;
    const unsigned int i = list[t];
    double esum = 0.0;
    if (i != NONE) {
        int ns[NNB];
        load_slots(nbr, t, ns);
        esum = list_energy(vols[v].L, vols[v].Cn, uc, c, t, i, uc[t], ns);
    }
    kf_block_publish<LB>(partials, w, esum, 0.0, 0ull);
}

// Block v: the energy partials of volume v in slot order -> energy[v].
template <class S>
__global__ __launch_bounds__(LB) void k_energy_fold_batch(const ActiveVol<S> *__restrict__ vols,
                                                          const BlockPartial *__restrict__ partials,
                                                          double *__restrict__ energy)
{
    const unsigned int b0 = vols[blockIdx.x].blk0, nb = vols[blockIdx.x].nblk;
    double s = 0.0, m = 0.0;
    unsigned long long c = 0;
    for (unsigned int b = threadIdx.x; b < nb; b += LB)
        s += partials[b0 + b].sum;
    kf_block_reduce3<LB>(s, m, c);
    if (threadIdx.x == 0)
        energy[blockIdx.x] = s;
}

// Each volume's final compact set back into its field.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_active_scatter_batch(const ActiveVol<S> *__restrict__ vols,
                                                                const unsigned int *__restrict__ blk_vol,
                                                                const float4 *__restrict__ u0, unsigned int set,
                                                                const DeformState *states,
                                                                const unsigned int *__restrict__ list,
                                                                unsigned int slots)
{
    const unsigned int t = blockIdx.x * (unsigned int)BLOCK + threadIdx.x;
    if (t >= slots)
        return;
    const unsigned int i = list[t];
    if (i == NONE)
        return;
    const unsigned int v = blk_vol[t / (unsigned int)LB];
    const float4 q = u0[(size_t)(states[v].iters & 1u) * set + t];
    float *u = vols[v].u;
    const unsigned int n32 = nvox32(vols[v].L.g);
    u[i] = q.x;
    u[n32 + i] = q.y;
    u[2u * n32 + i] = q.z;
}

__global__ void k_states_reset(DeformState *st, int n, double stop)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    DeformState s{};
    s.run[0] = s.run[1] = 1u;
    s.diverged = LLONG_MAX;
    s.div_slot[0] = s.div_slot[1] = LLONG_MAX;
    s.stop = stop;
    st[v] = s;
}

__global__ __launch_bounds__(BLOCK) void k_active_scatter(const unsigned int *__restrict__ list,
                                                          const float4 *__restrict__ u0, unsigned int set,
                                                          const DeformState *st, float *__restrict__ u,
                                                          unsigned int n32)
{
    const unsigned int t = blockIdx.x * (unsigned int)BLOCK + threadIdx.x;
    if (t >= st->list_n)
        return;
    const unsigned int i = list[t];
    const float4 v = u0[(size_t)(st->iters & 1u) * set + t];
    u[i] = v.x;
    u[(size_t)n32 + i] = v.y;
    u[2 * (size_t)n32 + i] = v.z;
}

template <class T>
kf_status grow(T *&p, size_t &cap, size_t need, hipStream_t s)
{
    if (need <= cap)
        return KF_OK;
    if (p) {
        KF_HIP(hipStreamSynchronize(s));
        KF_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
    }
    KF_HIP(hipMalloc((void **)&p, sizeof(T) * need));
    cap = need;
    return KF_OK;
}

}  // namespace

// Voxel indices, list entries and the slots of the 12 neighbours are 32-bit; plane offsets
// into the field are 64-bit, so any grid of fewer than 2^32 voxels can be listed.
bool kf_active_supported(const kf_volume *live) { return live->n < ((int64_t)1 << 32); }

void kf_active_release(kf_context *ctx)
{
    kf_active_ws &a = ctx->active;
    a.loop.drop();
    a.batch_loop.drop();
    (void)hipFree(a.mask);
    (void)hipFree(a.span_count);
    (void)hipFree(a.pos_of);
    (void)hipFree(a.list);
    (void)hipFree(a.u);
    (void)hipFree(a.nbr);
    (void)hipFree(a.bstates);
    (void)hipFree(a.benergy);
    (void)hipFree(a.blk_vol);
    (void)hipFree(a.vtab);
    (void)hipFree(a.ltab);
    (void)hipFree(a.span_vol);
    (void)hipFree(a.nb_vol);
    if (a.bstates_host)
        (void)hipHostFree(a.bstates_host);
    (void)hipFree(a.aux);
    if (a.aux_host)
        (void)hipHostFree(a.aux_host);
    (void)hipFree(a.wtmp);
    a = kf_active_ws{};
}

unsigned int kf_list_slots(unsigned int count, bool headroom)
{
    uint64_t c = count;
    if (headroom) {
        // an eighth more than the count, rounded up to a sixteenth of its power of two, so that
        // lists of similar length get the same launch shape (and the same captured loop)
        c += c / 8 + 1;
        int b = 63 - __builtin_clzll(c | 1ull);
        const uint64_t q = b > 14 ? (1ull << (b - 4)) : 1024ull;
        c = (c + q - 1) / q * q;
    }
    c = (c + SLOT_QUANTUM - 1) / SLOT_QUANTUM * SLOT_QUANTUM;
    if (c < SLOT_QUANTUM)
        c = SLOT_QUANTUM;
    return c > KF_LIST_MAX ? 0u : (unsigned int)c;
}

kf_status kf_active_count(const kf_volume *live, const float *u, const kf_params *p, bool u_zero, unsigned int slots,
                          bool need_finite)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    a.warp_vol = nullptr;
    KF_CHECK_ARG(kf_active_supported(live), "the active-set sweep needs N < 2^32");
    const AdjConsts c = kf_adj_consts(p);
    const unsigned int nspan = (unsigned int)((live->n + SPAN - 1) / SPAN);
    kf_status s = grow(a.mask, a.mask_cap, (size_t)nspan * WPB, ctx->stream);
    if (s == KF_OK)
        s = grow(a.span_count, a.span_cap, 2 * (size_t)nspan, ctx->stream);  // counts, then offsets
    if (s != KF_OK)
        return s;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        if (u_zero)
            hipLaunchKernelGGL((k_active_mark<S, false>), dim3(nspan), dim3(BLOCK), 0, ctx->stream, make_read<S>(live),
                               u, c.lo, c.hi, a.mask, a.span_count, &ctx->state->nonfinite);
        else
            hipLaunchKernelGGL((k_active_mark<S, true>), dim3(nspan), dim3(BLOCK), 0, ctx->stream, make_read<S>(live),
                               u, c.lo, c.hi, a.mask, a.span_count, &ctx->state->nonfinite);
    });
    KF_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_active_scan_sized, dim3(1), dim3(SCAN_BLOCK), 0, ctx->stream, a.span_count, nspan,
                       a.span_count + nspan, ctx->state, slots, need_finite ? 1 : 0);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_active_prepare(const kf_volume *live, const float *u, unsigned int slots, bool neighbours,
                            unsigned int reserve)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(slots > 0 && slots % SLOT_QUANTUM == 0, "list slots must be a positive multiple of the quantum");
    const unsigned int nspan = (unsigned int)((live->n + SPAN - 1) / SPAN);
    const unsigned int nwords = (unsigned int)((live->n + 63) / 64);
    kf_status s = grow(a.pos_of, a.pos_cap, (size_t)live->n, ctx->stream);
    if (s == KF_OK && (size_t)slots > a.cap) {
        // the list, two compact sets of u and the neighbour slots share one capacity
        const size_t want = reserve > slots ? reserve : slots;
        size_t c1 = a.cap, c2 = 2 * a.cap, c3 = 3 * a.cap;
        s = grow(a.list, c1, want, ctx->stream);
        if (s == KF_OK)
            s = grow(a.u, c2, 2 * want, ctx->stream);
        if (s == KF_OK)
            s = grow(a.nbr, c3, 3 * want, ctx->stream);
        a.cap = s == KF_OK ? want : 0;
    }
    if (s != KF_OK)
        return s;
    hipLaunchKernelGGL(k_active_compact_sized, dim3(nspan), dim3(WPB), 0, ctx->stream, a.mask, nwords, a.span_count,
                       a.span_count + nspan, (const DeformState *)ctx->state, a.list, a.pos_of);
    KF_LAUNCH_CHECK();
    if (neighbours) {
        hipLaunchKernelGGL(k_active_neighbours, dim3(slots / BLOCK), dim3(BLOCK), 0, ctx->stream, grid_of(live),
                           a.mask, a.pos_of, a.list, (const DeformState *)ctx->state, u, a.u, a.nbr);
        KF_LAUNCH_CHECK();
    }
    return KF_OK;
}

kf_status kf_launch_adj_list(const kf_volume *live, const kf_volume *can, const kf_params *p, int iter,
                             unsigned int slots, unsigned int expect, unsigned int *nb_out)
{
    // Per-iteration partials (max |du|, active count) feed only the stop rule and the last
    // iteration's statistics: without a stop rule only the last launch publishes them.
    const bool check = p->stop_max_update > 0.0;
    const bool publish = check || iter == p->max_iters - 1;
    const bool prev_published = check || iter == 0;  // launch iter - 1 published (never, for iter 0)
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    KF_CHECK_ARG(slots > 0 && slots % SLOT_QUANTUM == 0 && slots <= a.cap, "active list not prepared");
    KF_CHECK_ARG(expect <= slots && expect % SLOT_QUANTUM == 0, "expected list slots beyond the launch's");
    const AdjConsts c = kf_adj_consts(p);
    // nb work blocks over the list's slots (each publishes a partial) and the folding block.
    const unsigned int nb = slots / LB;
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    if (nb_out)
        *nb_out = nb;
    const int parity = iter & 1;
    const BlockPartial *prev = kf_partials(ctx, parity ^ 1);
    const unsigned int prev_nb = iter > 0 && prev_published ? nb : 0u;
    const float4 *uin = a.u + (size_t)parity * a.cap;
    float4 *uout = a.u + (size_t)(parity ^ 1) * a.cap;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_adj_list<S>), dim3(nb + 1), dim3(LB), 0, ctx->stream, make_read<S>(live),
                           make_read<S>(can), uin, uout, c, ctx->state, kf_partials(ctx, parity), parity, prev, prev_nb,
                           iter > 0 ? 1 : 0, publish ? 1 : 0, a.list, a.nbr, expect / LB);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_active_loop(const kf_volume *live, const kf_volume *can, const kf_params *p, unsigned int slots,
                         unsigned int expect)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    const int K = p->max_iters;
    KF_CHECK_ARG(slots > 0 && slots % SLOT_QUANTUM == 0 && slots <= a.cap, "active list not prepared");
    const unsigned int nb = slots / LB;
    kf_status s = kf_ensure_partials(ctx, nb);  // before any capture: the launches allocate nothing
    if (s != KF_OK || K <= 0)
        return s;
    const AdjConsts c = kf_adj_consts(p);
    char buf[320];
    std::snprintf(buf, sizeof(buf), "%d/%u/%u/%d/%p/%p/%p/%p/%p/%zu/%p/%zu/%p/%d/%d/%d/%d/%d", live->dtype, slots,
                  expect, K,
                  live->phi, can->phi, (void *)a.u, (void *)a.list, (void *)a.nbr, a.cap, (void *)ctx->partials,
                  ctx->partials_cap, (void *)ctx->state, live->dims[0], live->dims[1], live->dims[2],
                  (int)(live->scale() != can->scale()), (int)(p->stop_max_update > 0.0));
    std::string key(buf);
    key.append(reinterpret_cast<const char *>(&c), sizeof(c));
    const float sc[4] = {live->scale(), can->scale(), (float)live->oob(), (float)can->oob()};
    key.append(reinterpret_cast<const char *>(sc), sizeof(sc));
    return kf_run_captured(ctx, a.loop, key, [&]() -> kf_status {
        unsigned int nbo = 0;
        for (int k = 0; k < K; k++) {
            const kf_status r = kf_launch_adj_list(live, can, p, k, slots, expect, &nbo);
            if (r != KF_OK)
                return r;
        }
        return kf_launch_finalize(ctx, nbo, (K - 1) & 1);
    });
}

kf_status kf_active_energy(const kf_volume *live, const kf_volume *can, const kf_params *p, unsigned int slots,
                           unsigned int expect)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(slots > 0 && slots % SLOT_QUANTUM == 0 && slots <= a.cap, "active list not prepared");
    const AdjConsts c = kf_adj_consts(p);
    const unsigned int nb = slots / LB;
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_active_energy<S>), dim3(nb), dim3(LB), 0, ctx->stream, make_read<S>(live),
                           make_read<S>(can), a.u, (unsigned int)a.cap, ctx->state, c, kf_partials(ctx, 0), a.list,
                           a.nbr, expect / LB);
    });
    KF_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_list_energy_fold, dim3(1), dim3(1024), 0, ctx->stream, ctx->state,
                       (const BlockPartial *)kf_partials(ctx, 0), nb);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_active_scatter(const kf_volume *live, float *u, unsigned int slots)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(slots > 0 && slots % SLOT_QUANTUM == 0 && slots <= a.cap, "active list not prepared");
    hipLaunchKernelGGL(k_active_scatter, dim3(slots / BLOCK), dim3(BLOCK), 0, ctx->stream, a.list, a.u,
                       (unsigned int)a.cap, ctx->state, u, (unsigned int)live->n);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

namespace {

template <class S>
kf_status batch_active(kf_context *ctx, const kf_volume *const *live, const kf_volume *const *can,
                       kf_field *const *u, int count, const kf_params *p, kf_deform_stats *stats, bool force,
                       bool *used)
{
    kf_active_ws &a = ctx->active;
    hipStream_t st = ctx->stream;
    *used = false;
    a.warp_vol = nullptr;
    std::vector<size_t> moff(count), soff(count), poff(count);
    std::vector<unsigned int> nspan(count);
    size_t mtot = 0, stot = 0;
    int64_t nvox = 0;
    unsigned int sbt = 0;  // span blocks of all volumes
    for (int v = 0; v < count; v++) {
        nspan[v] = (unsigned int)((live[v]->n + SPAN - 1) / SPAN);
        moff[v] = mtot;
        mtot += (size_t)nspan[v] * WPB;
        soff[v] = stot;
        stot += 2 * (size_t)nspan[v];
        poff[v] = (size_t)nvox;
        nvox += live[v]->n;
        sbt += nspan[v];
    }
    kf_status s = grow(a.mask, a.mask_cap, mtot, st);
    if (s == KF_OK)
        s = grow(a.span_count, a.span_cap, stot, st);
    if (s == KF_OK)
        s = grow(a.pos_of, a.pos_cap, (size_t)nvox, st);
    if (s == KF_OK)
        s = grow(a.ltab, a.ltab_cap, sizeof(ListVol<S>) * count, st);
    if (s == KF_OK)
        s = grow(a.span_vol, a.span_vol_cap, (size_t)sbt, st);
    if (s == KF_OK && (size_t)count > a.bcap) {
        if (a.bstates_host) {
            KF_HIP(hipStreamSynchronize(st));
            (void)hipHostFree(a.bstates_host);
            a.bstates_host = nullptr;
        }
        size_t c1 = a.bcap, c2 = a.bcap;
        s = grow(a.bstates, c1, (size_t)count, st);
        if (s == KF_OK)
            s = grow(a.benergy, c2, (size_t)count, st);
        if (s == KF_OK && hipHostMalloc((void **)&a.bstates_host, (sizeof(DeformState) + sizeof(double)) * count,
                                        hipHostMallocDefault) != hipSuccess) {
            kf_set_error("pinned state allocation failed");
            s = KF_E_OOM;
        }
        a.bcap = s == KF_OK ? (size_t)count : 0;
    }
    if (s != KF_OK)
        return s;
    DeformState *sh = a.bstates_host;
    double *eh = reinterpret_cast<double *>(sh + count);
    const AdjConsts c = kf_adj_consts(p);
    hipLaunchKernelGGL(k_states_reset, dim3((count + 63) / 64), dim3(64), 0, st, a.bstates, count, p->stop_max_update);
    KF_LAUNCH_CHECK();
    // Mark and count every volume's candidates: one launch over all spans, one scan block per
    // volume.
    std::vector<ListVol<S>> lt(count);
    std::vector<unsigned int> sv(sbt);
    unsigned int sb0 = 0;
    for (int v = 0; v < count; v++) {
        ListVol<S> &L = lt[v];
        L.L = make_read<S>(live[v]);
        L.u = u[v]->u();
        L.mask = a.mask + moff[v];
        L.span_count = a.span_count + soff[v];
        L.pos_of = a.pos_of + poff[v];
        L.nonfinite = &a.bstates[v].nonfinite;
        L.listed = &a.bstates[v].listed;
        L.nspan = nspan[v];
        L.sb0 = sb0;
        L.base = L.cnt = L.nb0 = 0u;
        for (unsigned int b = 0; b < nspan[v]; b++)
            sv[sb0 + b] = (unsigned int)v;
        sb0 += nspan[v];
    }
    const ListVol<S> *dlt = reinterpret_cast<const ListVol<S> *>(a.ltab);
    KF_HIP(hipMemcpyAsync(a.ltab, lt.data(), sizeof(ListVol<S>) * count, hipMemcpyHostToDevice, st));
    KF_HIP(hipMemcpyAsync(a.span_vol, sv.data(), sizeof(unsigned int) * sbt, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_active_mark_batch<S>), dim3(sbt), dim3(BLOCK), 0, st, dlt, a.span_vol, c.lo, c.hi);
    KF_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_active_scan_batch<S>), dim3((unsigned int)count), dim3(SCAN_BLOCK), 0, st, dlt);
    KF_LAUNCH_CHECK();
    KF_HIP(hipMemcpyAsync(sh, a.bstates, sizeof(DeformState) * count, hipMemcpyDeviceToHost, st));
    KF_HIP(hipStreamSynchronize(st));
    std::vector<unsigned int> cnt(count), blk0(count), nblk(count);
    int64_t listed = 0;
    unsigned int nwb = 0;
    for (int v = 0; v < count; v++) {
        cnt[v] = sh[v].listed;
        listed += cnt[v];
        blk0[v] = nwb;
        nblk[v] = cnt[v] ? (cnt[v] + LB - 1) / LB : 1u;
        nwb += nblk[v];
    }
    if (!force && listed * 8 > nvox)
        return KF_OK;  // the dense batched sweep is cheaper
    *used = true;
    const size_t slots = (size_t)nwb * LB;
    if (slots > a.cap) {
        size_t c1 = a.cap, c2 = 2 * a.cap, c3 = 3 * a.cap;
        s = grow(a.list, c1, slots, st);
        if (s == KF_OK)
            s = grow(a.u, c2, 2 * slots, st);
        if (s == KF_OK)
            s = grow(a.nbr, c3, 3 * slots, st);
        a.cap = s == KF_OK ? slots : 0;
    }
    if (s == KF_OK)
        s = grow(a.blk_vol, a.blk_cap, (size_t)nwb, st);
    if (s == KF_OK)
        s = grow(a.vtab, a.vtab_cap, sizeof(ActiveVol<S>) * count, st);
    if (s == KF_OK)
        s = kf_ensure_partials(ctx, nwb);
    if (s != KF_OK)
        return s;
    std::vector<ActiveVol<S>> tab(count);
    std::vector<unsigned int> bv(nwb);
    for (int v = 0; v < count; v++) {
        tab[v].L = make_read<S>(live[v]);
        tab[v].Cn = make_read<S>(can[v]);
        tab[v].u = u[v]->u();
        tab[v].blk0 = blk0[v];
        tab[v].nblk = nblk[v];
        for (unsigned int b = 0; b < nblk[v]; b++)
            bv[blk0[v] + b] = (unsigned int)v;
    }
    const ActiveVol<S> *dtab = reinterpret_cast<const ActiveVol<S> *>(a.vtab);
    KF_HIP(hipMemcpyAsync(a.vtab, tab.data(), sizeof(ActiveVol<S>) * count, hipMemcpyHostToDevice, st));
    KF_HIP(hipMemcpyAsync(a.blk_vol, bv.data(), sizeof(unsigned int) * nwb, hipMemcpyHostToDevice, st));
    KF_HIP(hipMemsetAsync(a.list, 0xff, sizeof(unsigned int) * slots, st));
    // List them: one launch over all spans, one over all listed voxels.
    unsigned int nbt = 0;
    for (int v = 0; v < count; v++) {
        lt[v].base = blk0[v] * (unsigned int)LB;
        lt[v].cnt = cnt[v];
        lt[v].nb0 = nbt;
        nbt += (cnt[v] + BLOCK - 1) / BLOCK;
    }
    std::vector<unsigned int> nv(nbt);
    for (int v = 0; v < count; v++)
        for (unsigned int b = 0; b < (cnt[v] + BLOCK - 1) / BLOCK; b++)
            nv[lt[v].nb0 + b] = (unsigned int)v;
    if ((s = grow(a.nb_vol, a.nb_vol_cap, (size_t)(nbt ? nbt : 1u), st)) != KF_OK)
        return s;
    KF_HIP(hipMemcpyAsync(a.ltab, lt.data(), sizeof(ListVol<S>) * count, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_active_compact_batch<S>), dim3(sbt), dim3(WPB), 0, st, dlt, a.span_vol, a.list);
    KF_LAUNCH_CHECK();
    if (nbt > 0) {
        KF_HIP(hipMemcpyAsync(a.nb_vol, nv.data(), sizeof(unsigned int) * nbt, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_active_neighbours_batch<S>), dim3(nbt), dim3(BLOCK), 0, st, dlt, a.nb_vol, a.list, a.u,
                           a.nbr);
        KF_LAUNCH_CHECK();
    }
    // The iterations: launch k reads compact set k & 1 and writes the other one; per-volume
    // gates stop each volume on its own (stop rule or divergence).
    const bool check = p->stop_max_update > 0.0;
    const unsigned int set = (unsigned int)a.cap;
    int launched = 0;
    if (!check && p->max_iters > 0) {
        // Without a stop rule the loop needs no host wait: all launches and the last fold run
        // as one captured graph when the same batch shape comes back.
        const int K = p->max_iters;
        char buf[320];
        std::snprintf(buf, sizeof(buf), "%d/%d/%u/%d/%p/%p/%p/%u/%p/%p/%zu/%p/%p", (int)S::dtype, count, nwb, K,
                      (void *)dtab, (void *)a.blk_vol, (void *)a.u, set, (void *)a.bstates, (void *)ctx->partials,
                      ctx->partials_cap, (void *)a.list, (void *)a.nbr);
        std::string key(buf);
        key.append(reinterpret_cast<const char *>(&c), sizeof(c));
        s = kf_run_captured(ctx, a.batch_loop, key, [&]() -> kf_status {
            for (int k = 0; k < K; k++) {
                const int par = k & 1;
                hipLaunchKernelGGL((k_adj_list_batch<S>), dim3((unsigned int)count + nwb), dim3(LB), 0, st, dtab,
                                   count, a.blk_vol, a.u + (size_t)par * set, a.u + (size_t)(par ^ 1) * set, c,
                                   a.bstates, kf_partials(ctx, par), par, kf_partials(ctx, par ^ 1), k > 0 ? 1 : 0, 0,
                                   k == K - 1 ? 1 : 0, a.list, a.nbr);
                KF_LAUNCH_CHECK();
            }
            const int par = K & 1;
            hipLaunchKernelGGL((k_adj_list_batch<S>), dim3((unsigned int)count), dim3(LB), 0, st, dtab, count,
                               a.blk_vol, a.u, a.u, c, a.bstates, kf_partials(ctx, par), par,
                               kf_partials(ctx, par ^ 1), 1, 1, 0, a.list, a.nbr);
            KF_LAUNCH_CHECK();
            return KF_OK;
        });
        if (s != KF_OK)
            return s;
        launched = -1;  // folded already
    }
    for (int k = 0; launched >= 0 && k < p->max_iters; k++) {
        const int par = k & 1;
        const int publish = check || k == p->max_iters - 1;
        hipLaunchKernelGGL((k_adj_list_batch<S>), dim3((unsigned int)count + nwb), dim3(LB), 0, st, dtab, count,
                           a.blk_vol, a.u + (size_t)par * set, a.u + (size_t)(par ^ 1) * set, c, a.bstates,
                           kf_partials(ctx, par), par, kf_partials(ctx, par ^ 1), k > 0 ? 1 : 0, check ? 1 : 0,
                           publish, a.list, a.nbr);
        KF_LAUNCH_CHECK();
        launched = k + 1;
        if (check && launched % 64 == 0 && launched < p->max_iters) {
            KF_HIP(hipMemcpyAsync(sh, a.bstates, sizeof(DeformState) * count, hipMemcpyDeviceToHost, st));
            KF_HIP(hipStreamSynchronize(st));
            bool all = true;
            for (int v = 0; v < count; v++)
                all = all && sh[v].done;
            if (all)
                break;
        }
    }
    if (launched > 0) {
        // fold the last launch (fold blocks only)
        const int par = launched & 1;
        hipLaunchKernelGGL((k_adj_list_batch<S>), dim3((unsigned int)count), dim3(LB), 0, st, dtab, count, a.blk_vol,
                           a.u, a.u, c, a.bstates, kf_partials(ctx, par), par, kf_partials(ctx, par ^ 1), 1, 1, 0,
                           a.list, a.nbr);
        KF_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_active_scatter_batch<S>), dim3((unsigned int)((slots + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                       st, dtab, a.blk_vol, a.u, set, a.bstates, a.list, (unsigned int)slots);
    KF_LAUNCH_CHECK();
    if (stats) {
        hipLaunchKernelGGL((k_active_energy_batch<S>), dim3(nwb), dim3(LB), 0, st, dtab, a.blk_vol, a.u, set, c,
                           a.bstates, kf_partials(ctx, 0), a.list, a.nbr);
        KF_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_energy_fold_batch<S>), dim3((unsigned int)count), dim3(LB), 0, st, dtab,
                           kf_partials(ctx, 0), a.benergy);
        KF_LAUNCH_CHECK();
        KF_HIP(hipMemcpyAsync(eh, a.benergy, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    }
    KF_HIP(hipMemcpyAsync(sh, a.bstates, sizeof(DeformState) * count, hipMemcpyDeviceToHost, st));
    KF_HIP(hipStreamSynchronize(st));
    kf_status worst = KF_OK;
    std::string msg;
    for (int v = 0; v < count; v++) {
        const DeformState &h = sh[v];
        if (h.diverged != LLONG_MAX) {
            worst = KF_E_DIVERGED;
            if (msg.empty())
                msg = "volume " + std::to_string(v) + ": deformation field diverged at voxel " +
                      std::to_string((long long)h.diverged);
        }
        if (stats) {
            kf_deform_stats &o = stats[v];
            std::memset(&o, 0, sizeof(o));
            o.iters = (int32_t)h.iters;
            o.sweep = KF_SWEEP_ACTIVE;
            o.diverged_index = h.diverged == LLONG_MAX ? -1 : (int64_t)h.diverged;
            o.max_update = h.last_max;
            o.energy = h.diverged == LLONG_MAX ? eh[v] : 0.0;
            o.active_voxels = h.last_active;
            o.listed_voxels = cnt[v];
        }
    }
    if (worst != KF_OK)
        kf_set_error(msg);
    return worst;
}

}  // namespace

kf_status kf_batch_active(kf_context *ctx, const kf_volume *const *live, const kf_volume *const *can,
                          kf_field *const *u, int count, const kf_params *p, kf_deform_stats *stats, bool force,
                          bool *used)
{
    KF_DISPATCH_DTYPE(live[0]->dtype, S, return batch_active<S>(ctx, live, can, u, count, p, stats, force, used));
    return KF_OK;
}

kf_status kf_list_moved(const kf_volume *v, const float *u, unsigned int *count, bool *nonfinite)
{
    kf_context *ctx = v->ctx;
    kf_active_ws &a = ctx->active;
    a.warp_vol = nullptr;
    hipStream_t st = ctx->stream;
    KF_CHECK_ARG(kf_active_supported(v), "listing needs N < 2^32");
    const unsigned int nspan = (unsigned int)((v->n + SPAN - 1) / SPAN);
    const unsigned int nwords = (unsigned int)((v->n + 63) / 64);
    kf_status s = grow(a.mask, a.mask_cap, (size_t)nspan * WPB, st);
    if (s == KF_OK)
        s = grow(a.span_count, a.span_cap, 2 * (size_t)nspan, st);
    if (s == KF_OK)
        s = grow(a.pos_of, a.pos_cap, (size_t)v->n, st);
    if (s == KF_OK && !a.aux) {
        KF_HIP(hipMalloc((void **)&a.aux, 4 * sizeof(unsigned int)));
        if (hipHostMalloc((void **)&a.aux_host, 4 * sizeof(unsigned int), hipHostMallocDefault) != hipSuccess) {
            kf_set_error("pinned counter allocation failed");
            s = KF_E_OOM;
        }
    }
    if (s != KF_OK)
        return s;
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, st, a.aux, 4);
    KF_LAUNCH_CHECK();
    KF_DISPATCH_DTYPE(v->dtype, S, {
        hipLaunchKernelGGL((k_moved_mark<S>), dim3(nspan), dim3(BLOCK), 0, st, (const typename S::T *)v->phi,
                           v->scale(), u, (unsigned int)v->n, a.mask, a.span_count, a.aux + 1);
    });
    KF_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_active_scan, dim3(1), dim3(SCAN_BLOCK), 0, st, a.span_count, nspan, a.span_count + nspan,
                       a.aux);
    KF_LAUNCH_CHECK();
    KF_HIP(hipMemcpyAsync(a.aux_host, a.aux, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    KF_HIP(hipStreamSynchronize(st));
    *count = a.aux_host[0];
    *nonfinite = a.aux_host[1] != 0u;
    if (*nonfinite || *count == 0)
        return KF_OK;
    if ((size_t)*count > a.cap) {
        size_t c1 = a.cap, c2 = 2 * a.cap, c3 = 3 * a.cap;
        s = grow(a.list, c1, *count, st);
        if (s == KF_OK)
            s = grow(a.u, c2, 2 * (size_t)*count, st);
        if (s == KF_OK)
            s = grow(a.nbr, c3, 3 * (size_t)*count, st);
        a.cap = s == KF_OK ? *count : 0;
        if (s != KF_OK)
            return s;
    }
    hipLaunchKernelGGL(k_active_compact, dim3(nspan), dim3(WPB), 0, st, a.mask, nwords, a.span_count,
                       a.span_count + nspan, 0u, a.list, a.pos_of);
    KF_LAUNCH_CHECK();
    return KF_OK;
}
