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
//    per span of 16384 voxels; k_active_scan turns the counts into span offsets and the total.
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

namespace {

constexpr int LB = 64;                 // threads per block of the list step (one wavefront)
constexpr int SPAN = BLOCK * 16;       // voxels per block of the marking and listing passes
constexpr int WPB = SPAN / 64;         // 64-voxel mask words per such block (one per thread)
constexpr int NNB = 12;                // neighbours of the Killing term: 6 faces, 6 edges
constexpr int SCAN_BLOCK = 1024;

// Neighbour k of voxel (x, y, z): faces +x -x +y -y +z -z, then the edge neighbours of Nbhd::dgl
// in the order dgl[0][1], dgl[0][2], dgl[1][0], dgl[1][2], dgl[2][0], dgl[2][1].
__constant__ int c_nb_off[NNB][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                                     {1, -1, 0}, {1, 0, -1}, {-1, 1, 0}, {0, 1, -1}, {-1, 0, 1}, {0, -1, 1}};

template <class S>
__global__ __launch_bounds__(BLOCK) void k_active_mark(PhiRead<S> L, const float *__restrict__ u, float lo, float hi,
                                                       unsigned long long *__restrict__ mask,
                                                       unsigned int *__restrict__ span_count)
{
    const unsigned int n32 = nvox32(L.g);
    const unsigned int base = blockIdx.x * (unsigned int)SPAN;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int cnt = 0;
#pragma unroll 4
    for (int k = 0; k < SPAN / BLOCK; k++) {
        const unsigned int w0 = base + (unsigned int)(k * BLOCK + w * 64);
        const unsigned int i = w0 + (unsigned int)lane;
        bool cand = false;
        if (i < n32) {
            const float u0 = kf_load_stream(u + i), u1 = kf_load_stream(u + n32 + i);
            const float u2 = kf_load_stream(u + 2u * n32 + i);
            const float live = S::load(L.p, (int64_t)i, L.scale);
            cand = (__float_as_uint(u0) | __float_as_uint(u1) | __float_as_uint(u2)) != 0u ||
                   !(live > hi || live < lo);
        }
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

// Exclusive scan of the span counts (one block); the total goes to st->listed.
__global__ __launch_bounds__(SCAN_BLOCK) void k_active_scan(const unsigned int *__restrict__ span_count,
                                                            unsigned int nspan, unsigned int *__restrict__ span_off,
                                                            DeformState *st)
{
    const unsigned int per = (nspan + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const unsigned int b0 = threadIdx.x * per;
    unsigned int sum = 0;
    for (unsigned int k = 0; k < per; k++)
        if (b0 + k < nspan)
            sum += span_count[b0 + k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned int v = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int t = __shfl_up(v, o, 64);
        if (lane >= o)
            v += t;
    }
    __shared__ unsigned int s_w[SCAN_BLOCK / 64];
    if (lane == 63)
        s_w[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int r = 0;
        for (int k = 0; k < SCAN_BLOCK / 64; k++) {
            const unsigned int t = s_w[k];
            s_w[k] = r;
            r += t;
        }
    }
    __syncthreads();
    unsigned int run = s_w[w] + v - sum;
    for (unsigned int k = 0; k < per; k++)
        if (b0 + k < nspan) {
            span_off[b0 + k] = run;
            run += span_count[b0 + k];
        }
    if (threadIdx.x == SCAN_BLOCK - 1)
        st->listed = run;
}

// Thread k of block b lists the voxels of mask word b WPB + k, in voxel order.
__global__ __launch_bounds__(WPB) void k_active_compact(const unsigned long long *__restrict__ mask, unsigned int nwords,
                                                        const unsigned int *__restrict__ span_count,
                                                        const unsigned int *__restrict__ span_off,
                                                        unsigned int *__restrict__ list, unsigned int *__restrict__ pos_of)
{
    if (span_count[blockIdx.x] == 0u)
        return;
    const unsigned int wd = blockIdx.x * (unsigned int)WPB + threadIdx.x;
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
    unsigned int pos = span_off[blockIdx.x] + wbase + v - c;
    while (m) {
        const unsigned int i = wd * 64u + (unsigned int)__builtin_ctzll(m);
        m &= m - 1ull;
        list[pos] = i;
        pos_of[i] = pos;
        pos++;
    }
}

// Slot t: u of the listed voxel (compact set 0) and the list slots of its 12 neighbours.
__global__ __launch_bounds__(BLOCK) void k_active_neighbours(GridDims g, const unsigned long long *__restrict__ mask,
                                                             const unsigned int *__restrict__ pos_of,
                                                             const unsigned int *__restrict__ list, unsigned int count,
                                                             const float *__restrict__ u, float4 *__restrict__ uc,
                                                             int4 *__restrict__ nbr)
{
    const unsigned int t = blockIdx.x * (unsigned int)BLOCK + threadIdx.x;
    if (t >= count)
        return;
    const unsigned int i = list[t];
    const unsigned int n32 = nvox32(g);
    uc[t] = make_float4(u[i], u[n32 + i], u[2u * n32 + i], 0.0f);
    const unsigned int nxy = (unsigned int)g.nx * (unsigned int)g.ny;
    const int z = (int)(i / nxy), r = (int)(i - (unsigned int)z * nxy);
    const int y = r / g.nx, x = r - y * g.nx;
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

// Fold of step launch k - 1 by block 0 of launch k: kf_fold_launch's result, with every load
// (the gate flag, the state fields the fold reads and the partials) issued before any of them
// is needed, so the fold costs one round trip.  Partials are read even when the gate is
// closed; their values are then ignored.
template <int NT>
__device__ inline void list_fold(DeformState *st, const BlockPartial *prev, unsigned int prev_nb, int parity)
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
        if (go && prev_nb > 0 && !done) {
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
    const unsigned int nxy = (unsigned int)g.nx * (unsigned int)g.ny;
    z = (int)(i / nxy);
    const int r = (int)(i - (unsigned int)z * nxy);
    y = r / g.nx;
    x = r - y * g.nx;
}

// One Jacobi iteration over the list.  Block 0 only folds the previous launch and arms the
// next one (step_gate's first-block duty), beside the stencils rather than in front of one.
// A work block loads its list entries, own u and neighbour slots together with the gate flag
// (one round trip), then the live cell and the neighbours' u (a second one).
template <class S>
__global__ __launch_bounds__(LB) void k_adj_list(PhiRead<S> L, PhiRead<S> Cn, const float4 *__restrict__ uin,
                                                 float4 *__restrict__ uout, AdjConsts c, DeformState *st,
                                                 BlockPartial *partials, int parity, const BlockPartial *prev,
                                                 unsigned int prev_nb, const unsigned int *__restrict__ list,
                                                 const int4 *__restrict__ nbr, unsigned int count)
{
    if (blockIdx.x == 0) {
        list_fold<LB>(st, prev, prev_nb, parity);
        return;
    }
    const unsigned int nw = gridDim.x - 1u, w = blockIdx.x - 1u;
    // Consecutive runs of the list (neighbouring voxels) go to one XCD.
    const unsigned int t = kf_xcd_tile(w, nw) * (unsigned int)LB + threadIdx.x;
    const bool mine = t < count;
    unsigned int i = 0u;
    int ns[NNB];
    float4 own = make_float4(0.f, 0.f, 0.f, 0.f);
    if (mine) {
        i = list[t];
        own = uin[t];
        load_slots(nbr, t, ns);
    }
    if (st->run[parity] == 0u)
        return;
    double mx = 0.0;
    unsigned long long act = 0;
    if (mine) {
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
        adj_update(c, o, nb.uc, (int64_t)i, &st->div_slot[parity], mx, n);
        uout[t] = make_float4(n[0], n[1], n[2], 0.0f);
    }
    kf_block_publish<LB>(partials, w, 0.0, mx, act);
}

// Energy of field `uc` (see the file header of k_deform_adj.hip for E): slot t adds the term
// of its listed voxel j and of each unlisted voxel i = j - e_b (in the grid) for which j is
// the first listed voxel among i + e_0, i + e_1, i + e_2 -- so every unlisted voxel next to a
// listed one is counted once.  An unlisted voxel holds u = +0 and phi_live[i] outside the band
// (it is inactive), so only its Killing term is nonzero, and that reads u at i + e_b only.
template <class S>
__global__ __launch_bounds__(LB) void k_active_energy(PhiRead<S> L, PhiRead<S> Cn, const float4 *__restrict__ uc,
                                                      AdjConsts c, BlockPartial *partials,
                                                      const unsigned int *__restrict__ list,
                                                      const int4 *__restrict__ nbr, unsigned int count)
{
    const unsigned int t = kf_xcd_tile(blockIdx.x, gridDim.x) * (unsigned int)LB + threadIdx.x;
    double esum = 0.0;
    if (t < count) {
        const unsigned int i = list[t];
        const float4 own = uc[t];
        int ns[NNB];
        load_slots(nbr, t, ns);
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
        esum += (double)o.e;
        // unlisted neighbours on the low side: i - e_b, owned when no i - e_b + e_b' (b' < b)
        // is listed.  Slots: faces -x -y -z = 1, 3, 5; i - e_b + e_b' are edge slots
        // (b, b') = (1, 0): 6, (2, 0): 7, (2, 1): 9; i - e_b + e_b'' (b'' > b): (0, 1): 8,
        // (0, 2): 10, (1, 2): 11.
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
    }
    kf_block_publish<LB>(partials, blockIdx.x, esum, 0.0, 0ull);
}

__global__ __launch_bounds__(BLOCK) void k_active_scatter(const unsigned int *__restrict__ list,
                                                          const float4 *__restrict__ uc, unsigned int count,
                                                          float *__restrict__ u, unsigned int n32)
{
    const unsigned int t = blockIdx.x * (unsigned int)BLOCK + threadIdx.x;
    if (t >= count)
        return;
    const unsigned int i = list[t];
    const float4 v = uc[t];
    u[i] = v.x;
    u[n32 + i] = v.y;
    u[2u * n32 + i] = v.z;
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

bool kf_active_supported(const kf_volume *live) { return 3 * live->n < ((int64_t)1 << 30); }

void kf_active_release(kf_context *ctx)
{
    kf_active_ws &a = ctx->active;
    (void)hipFree(a.mask);
    (void)hipFree(a.span_count);
    (void)hipFree(a.pos_of);
    (void)hipFree(a.list);
    (void)hipFree(a.u);
    (void)hipFree(a.nbr);
    a = kf_active_ws{};
}

kf_status kf_active_count(const kf_volume *live, const float *u, const kf_params *p)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(kf_active_supported(live), "the active-set sweep needs 3 N < 2^30");
    const AdjConsts c = kf_adj_consts(p);
    const unsigned int nspan = (unsigned int)((live->n + SPAN - 1) / SPAN);
    kf_status s = grow(a.mask, a.mask_cap, (size_t)nspan * WPB, ctx->stream);
    if (s == KF_OK)
        s = grow(a.span_count, a.span_cap, 2 * (size_t)nspan, ctx->stream);  // counts, then offsets
    if (s != KF_OK)
        return s;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_active_mark<S>), dim3(nspan), dim3(BLOCK), 0, ctx->stream, make_read<S>(live), u, c.lo,
                           c.hi, a.mask, a.span_count);
    });
    KF_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_active_scan, dim3(1), dim3(SCAN_BLOCK), 0, ctx->stream, a.span_count, nspan,
                       a.span_count + nspan, ctx->state);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_active_prepare(const kf_volume *live, const float *u, unsigned int count)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    const unsigned int nspan = (unsigned int)((live->n + SPAN - 1) / SPAN);
    const unsigned int nwords = (unsigned int)((live->n + 63) / 64);
    kf_status s = grow(a.pos_of, a.pos_cap, (size_t)live->n, ctx->stream);
    if (s == KF_OK && (size_t)count > a.cap) {
        // the list, two compact sets of u and the neighbour slots share one capacity
        size_t c1 = a.cap, c2 = 2 * a.cap, c3 = 3 * a.cap;
        s = grow(a.list, c1, count, ctx->stream);
        if (s == KF_OK)
            s = grow(a.u, c2, 2 * (size_t)count, ctx->stream);
        if (s == KF_OK)
            s = grow(a.nbr, c3, 3 * (size_t)count, ctx->stream);
        a.cap = s == KF_OK ? count : 0;
    }
    if (s != KF_OK)
        return s;
    hipLaunchKernelGGL(k_active_compact, dim3(nspan), dim3(WPB), 0, ctx->stream, a.mask, nwords, a.span_count,
                       a.span_count + nspan, a.list, a.pos_of);
    KF_LAUNCH_CHECK();
    if (count > 0) {
        hipLaunchKernelGGL(k_active_neighbours, dim3((count + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream,
                           grid_of(live), a.mask, a.pos_of, a.list, count, u, a.u, a.nbr);
        KF_LAUNCH_CHECK();
    }
    return KF_OK;
}

kf_status kf_launch_adj_list(const kf_volume *live, const kf_volume *can, const kf_params *p, int iter,
                             unsigned int count, unsigned int *nb_out)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    KF_CHECK_ARG(count <= a.cap, "active list not prepared");
    const AdjConsts c = kf_adj_consts(p);
    // nb work blocks (at least one, publishing a partial) and the folding block.
    const unsigned int nb = count ? (count + LB - 1) / LB : 1u;
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    if (nb_out)
        *nb_out = nb;
    const int parity = iter & 1;
    const BlockPartial *prev = kf_partials(ctx, parity ^ 1);
    const unsigned int prev_nb = iter > 0 ? nb : 0u;
    const float4 *uin = a.u + (size_t)parity * a.cap;
    float4 *uout = a.u + (size_t)(parity ^ 1) * a.cap;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_adj_list<S>), dim3(nb + 1), dim3(LB), 0, ctx->stream, make_read<S>(live),
                           make_read<S>(can), uin, uout, c, ctx->state, kf_partials(ctx, parity), parity, prev, prev_nb,
                           a.list, a.nbr, count);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_active_energy(const kf_volume *live, const kf_volume *can, const kf_params *p, unsigned int count,
                           unsigned int iters)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    KF_CHECK_ARG(count <= a.cap, "active list not prepared");
    const AdjConsts c = kf_adj_consts(p);
    const unsigned int nb = count ? (count + LB - 1) / LB : 1u;
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s == KF_OK)
        s = kf_state_reset(ctx, 0.0);
    if (s != KF_OK)
        return s;
    const float4 *uc = a.u + (size_t)(iters & 1u) * a.cap;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        hipLaunchKernelGGL((k_active_energy<S>), dim3(nb), dim3(LB), 0, ctx->stream, make_read<S>(live),
                           make_read<S>(can), uc, c, kf_partials(ctx, 0), a.list, a.nbr, count);
    });
    KF_LAUNCH_CHECK();
    return kf_launch_finalize(ctx, nb, 0);
}

kf_status kf_active_scatter(const kf_volume *live, float *u, unsigned int count, unsigned int iters)
{
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    if (count == 0)
        return KF_OK;
    const float4 *uc = a.u + (size_t)(iters & 1u) * a.cap;
    hipLaunchKernelGGL(k_active_scatter, dim3((count + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, a.list, uc,
                       count, u, (unsigned int)live->n);
    KF_LAUNCH_CHECK();
    return KF_OK;
}
