// k_deform_batch.hip -- adjoint deform of independent volumes of uneven size.
//
// Every iteration is one persistent launch over the tiles of all volumes.  Block b first
// folds the previous launch of volumes b, b + G, ... (G = grid size) and arms their gates,
// then strides over the batch's tile space.  Each volume keeps its own gate, stop rule,
// divergence slot and partial range, so volumes stop independently, exactly as kf_deform
// (kf_api.hip) stops one volume; the per-voxel arithmetic is kf_adjoint.hpp's, so results
// equal kf_deform's bit for bit.
//
// A tile is BLOCK VPT consecutive voxels of one volume's flat index; thread t owns voxels
// tile0 + 256 k + t (k < VPT), so every load and store of a wavefront is 256 contiguous bytes and
// rows of any length leave no lane idle.  The launch sequence
// (state reset, K steps, last fold) is captured once into a hipGraph and replayed while the
// batch keeps its shape; the tile shape and grid size are the fastest of the candidates
// timed on the batch itself (kf_batch_autotune).
#include "kf_adjoint.hpp"

#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace {

template <class S>
struct BatchVol {
    PhiRead<S> L, Cn;
    float *u[2];          // launch k reads u[k & 1] and writes u[(k + 1) & 1]
    unsigned int tile0;   // first tile of the volume in the batch's tile space
    unsigned int ntiles;
    int vec4;             // N is a multiple of 4: tiles stream as aligned float4 groups
    int pad;
};

// One tile of a Jacobi step: stream u and phi_live, store every voxel's u_j, then run the
// stencil for the voxels that are active or already moved, compacted over the wavefront.
template <class S, int VPT>
__device__ inline void batch_tile(const BatchVol<S> &V, unsigned int t, const AdjConsts &c, long long *div,
                                  int parity, double &mx, unsigned long long &act)
{
    const GridDims g = V.L.g;
    const unsigned int n32 = nvox32(g);
    const unsigned int base = t * (unsigned int)(BLOCK * VPT);
    const float *__restrict__ u_in = V.u[parity];
    float *__restrict__ u_out = V.u[parity ^ 1];

    float cu[VPT][3], live[VPT];
    bool ok[VPT];
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const unsigned int i = base + (unsigned int)k * BLOCK + threadIdx.x;
        ok[k] = i < n32;
        const unsigned int ii = ok[k] ? i : 0u;
        cu[k][0] = ok[k] ? kf_load_stream(u_in + ii) : 0.f;
        cu[k][1] = ok[k] ? kf_load_stream(u_in + n32 + ii) : 0.f;
        cu[k][2] = ok[k] ? kf_load_stream(u_in + 2u * n32 + ii) : 0.f;
        live[k] = ok[k] ? S::load(V.L.p, (int64_t)ii, V.L.scale) : 0.f;
    }
    unsigned int need = 0u;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        if (!ok[k])
            continue;
        const unsigned int i = base + (unsigned int)k * BLOCK + threadIdx.x;
        u_out[i] = cu[k][0];
        u_out[n32 + i] = cu[k][1];
        u_out[2u * n32 + i] = cu[k][2];
        const bool moved = ((__float_as_uint(cu[k][0]) | __float_as_uint(cu[k][1]) | __float_as_uint(cu[k][2])) &
                            0x7fffffffu) != 0u;
        const bool in_band = !(live[k] > c.hi || live[k] < c.lo);
        if (moved || in_band)
            need |= 1u << k;
    }
    // Phase 2 as in k_adj_flat: one job list per wavefront; a job's stores may hit elements
    // another lane stored above, and a wavefront's vector memory operations are performed in
    // order.
    unsigned long long bal[VPT];
    int total = 0;
#pragma unroll
    for (int b = 0; b < VPT; b++) {
        bal[b] = __ballot((need >> b) & 1u);
        total += __popcll(bal[b]);
    }
    double esum = 0.0;
    const unsigned int wave0 = base + (threadIdx.x & ~63u);
    for (int j = threadIdx.x & 63; j < total; j += 64) {
        int bit, src;
        kf_job_decode<VPT>(bal, j, bit, src);
        const unsigned int i = wave0 + (unsigned int)bit * BLOCK + (unsigned int)src;
        int x, y, z;
        g.xyz32(i, x, y, z);
        adj_job<S, KF_MODE_STEP>(V.L, V.Cn, u_in, u_out, c, div, nullptr, x, y, z, mx, esum, act);
    }
}

// The same tile streamed as float4 groups like k_adj_flat: thread t owns groups
// base + 1024 q + 4 t (q < G), so a tile still holds BLOCK 4 G consecutive voxels.  UNAL: N is
// not a multiple of 4, so the u planes a n (a = 1, 2) start off a 16-byte boundary; their groups
// are read and written as 4-byte-aligned 16-byte accesses, and the one group that runs past N
// goes element by element.
template <class S, int G, bool UNAL>
__device__ inline void batch_tile_v4(const BatchVol<S> &V, unsigned int t, const AdjConsts &c, long long *div,
                                     int parity, double &mx, unsigned long long &act)
{
    constexpr int VEC = 4;
    const GridDims g = V.L.g;
    const unsigned int n32 = nvox32(g);
    const unsigned int base = t * (unsigned int)(BLOCK * VEC * G);
    const float *__restrict__ u_in = V.u[parity];
    float *__restrict__ u_out = V.u[parity ^ 1];
    unsigned int i0[G];
    bool ok[G];
    float4 C[G][3];
    float live4[G][VEC];
    unsigned int nin[G];  // voxels of the group inside the grid
#pragma unroll
    for (int q = 0; q < G; q++) {
        i0[q] = base + (unsigned int)q * (BLOCK * VEC) + threadIdx.x * VEC;
        ok[q] = i0[q] < n32;
        nin[q] = !ok[q] ? 0u : (UNAL && n32 - i0[q] < (unsigned int)VEC ? n32 - i0[q] : (unsigned int)VEC);
#pragma unroll
        for (int k = 0; k < VEC; k++)
            live4[q][k] = 0.f;
        if (!UNAL || nin[q] == (unsigned int)VEC) {
#pragma unroll
            for (int a = 0; a < 3; a++)
                C[q][a] = !ok[q] ? make_float4(0.f, 0.f, 0.f, 0.f)
                          : UNAL ? kf_load_stream4u(u_in + ((unsigned int)a * n32 + i0[q]))
                                 : kf_load_stream4(u_in + ((unsigned int)a * n32 + i0[q]));
            if (ok[q])
                S::load4(V.L.p, (int64_t)i0[q], V.L.scale, live4[q]);
        } else {
            float tv[3][VEC];
#pragma unroll
            for (int k = 0; k < VEC; k++) {
#pragma unroll
                for (int a = 0; a < 3; a++)
                    tv[a][k] = (unsigned int)k < nin[q] ? u_in[(unsigned int)a * n32 + i0[q] + (unsigned int)k] : 0.f;
                if ((unsigned int)k < nin[q])
                    live4[q][k] = S::load(V.L.p, (int64_t)(i0[q] + (unsigned int)k), V.L.scale);
            }
#pragma unroll
            for (int a = 0; a < 3; a++)
                C[q][a] = make_float4(tv[a][0], tv[a][1], tv[a][2], tv[a][3]);
        }
    }
    unsigned int need = 0u;
#pragma unroll
    for (int q = 0; q < G; q++) {
        if (!ok[q])
            continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float *dst = u_out + ((unsigned int)a * n32 + i0[q]);
            if (!UNAL) {
                *reinterpret_cast<float4 *>(dst) = C[q][a];
            } else if (nin[q] == (unsigned int)VEC) {
                *reinterpret_cast<kf_f4u *>(dst) = kf_f4u{C[q][a].x, C[q][a].y, C[q][a].z, C[q][a].w};
            } else {
                const float e[VEC] = {C[q][a].x, C[q][a].y, C[q][a].z, C[q][a].w};
#pragma unroll
                for (int k = 0; k < VEC; k++)
                    if ((unsigned int)k < nin[q])
                        dst[k] = e[k];
            }
        }
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            if (UNAL && (unsigned int)k >= nin[q])
                continue;
            const float c0 = k == 0 ? C[q][0].x : k == 1 ? C[q][0].y : k == 2 ? C[q][0].z : C[q][0].w;
            const float c1 = k == 0 ? C[q][1].x : k == 1 ? C[q][1].y : k == 2 ? C[q][1].z : C[q][1].w;
            const float c2 = k == 0 ? C[q][2].x : k == 1 ? C[q][2].y : k == 2 ? C[q][2].z : C[q][2].w;
            const bool moved = ((__float_as_uint(c0) | __float_as_uint(c1) | __float_as_uint(c2)) & 0x7fffffffu) != 0u;
            const bool in_band = !(live4[q][k] > c.hi || live4[q][k] < c.lo);
            if (moved || in_band)
                need |= 1u << (q * VEC + k);
        }
    }
    unsigned long long bal[G * VEC];
    int total = 0;
#pragma unroll
    for (int b = 0; b < G * VEC; b++) {
        bal[b] = __ballot((need >> b) & 1u);
        total += __popcll(bal[b]);
    }
    double esum = 0.0;
    const unsigned int wave0 = base + (threadIdx.x & ~63u) * VEC;
    for (int j = threadIdx.x & 63; j < total; j += 64) {
        int bit, src;
        kf_job_decode<G * VEC>(bal, j, bit, src);
        const unsigned int i = wave0 + (unsigned int)(bit / VEC) * (BLOCK * VEC) + (unsigned int)src * VEC +
                               (unsigned int)(bit % VEC);
        int x, y, z;
        g.xyz32(i, x, y, z);
        adj_job<S, KF_MODE_STEP>(V.L, V.Cn, u_in, u_out, c, div, nullptr, x, y, z, mx, esum, act);
    }
}

// Step launch `iter` of the batch.  parts holds two sets of cap slots (launch parity); the
// tile t of a launch publishes into slot t of its set.
template <class S, int VPT>
__global__ __launch_bounds__(BLOCK) void k_adj_batch(const BatchVol<S> *__restrict__ vols, int count,
                                                     unsigned int total, AdjConsts c, DeformState *states,
                                                     BlockPartial *parts, unsigned int cap, int iter)
{
    const int parity = iter & 1;
    // Gates: volume v runs launch k iff run[k & 1]; the fold of launch k - 1 arms run[(k + 1) & 1],
    // which no block of this launch reads.
    for (int v = blockIdx.x; v < count; v += gridDim.x) {
        DeformState *st = states + v;
        const bool go = st->run[parity] != 0u;
        if (go && iter > 0) {
            kf_fold_launch<BLOCK>(st, parts + (size_t)(parity ^ 1) * cap + vols[v].tile0, vols[v].ntiles, parity ^ 1);
            __syncthreads();  // the LDS of the fold is reused below
        }
        if (threadIdx.x == 0)
            st->run[parity ^ 1] = (go && !st->done) ? 1u : 0u;
    }
    BlockPartial *mine = parts + (size_t)parity * cap;
    // Odd launches visit the tiles backwards, so they start on the lines the previous launch
    // wrote last (still in the last-level cache).
    for (unsigned int tt = blockIdx.x; tt < total; tt += gridDim.x) {
        // tiles go to the XCDs (dispatch index mod 8) in chunks of 64, dealt round-robin
        const unsigned int tx = kf_xcd_chunk_tile<64>(tt, total);
        const unsigned int t = parity ? total - 1u - tx : tx;
        int lo = 0, hi = count - 1;  // volume of tile t (tile0 ascending)
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (vols[mid].tile0 <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        if (states[lo].run[parity] == 0u)
            continue;
        double mx = 0.0;
        unsigned long long act = 0;
        if (VPT >= 4 && vols[lo].vec4)
            batch_tile_v4<S, (VPT >= 4 ? VPT / 4 : 1), false>(vols[lo], t - vols[lo].tile0, c,
                                                              &states[lo].div_slot[parity], parity, mx, act);
        else if (VPT >= 4)
            batch_tile_v4<S, (VPT >= 4 ? VPT / 4 : 1), true>(vols[lo], t - vols[lo].tile0, c,
                                                             &states[lo].div_slot[parity], parity, mx, act);
        else
            batch_tile<S, VPT>(vols[lo], t - vols[lo].tile0, c, &states[lo].div_slot[parity], parity, mx, act);
        kf_block_publish<BLOCK>(mine, t, 0.0, mx, act);
        __syncthreads();  // thread 0 reads the publish LDS before the next tile writes it
    }
}

// Folds each volume's last launch (one block per volume) if that launch ran for it.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_batch_finish(const BatchVol<S> *__restrict__ vols, DeformState *states,
                                                        const BlockPartial *parts, unsigned int cap, int last)
{
    const int v = blockIdx.x;
    DeformState *st = states + v;
    if (st->run[last & 1] == 0u)
        return;
    kf_fold_launch<BLOCK>(st, parts + (size_t)(last & 1) * cap + vols[v].tile0, vols[v].ntiles, last & 1);
}

__global__ void k_batch_reset(DeformState *states, int count, double stop)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < count) {
        DeformState s{};
        s.run[0] = 1u;
        s.run[1] = 1u;
        s.diverged = LLONG_MAX;
        s.div_slot[0] = LLONG_MAX;
        s.div_slot[1] = LLONG_MAX;
        s.stop = stop;
        states[v] = s;
    }
}

struct BatchShape {
    int vpt = 4;
    unsigned int blocks = 0;  // 0: one block per tile
    int candidates = 0;
    double us = 0.0;
};

constexpr int VPTS[4] = {1, 2, 4, 8};
constexpr int GRID_MULT[4] = {0, 2, 4, 8};  // blocks = mult x CUs (0: one block per tile)

using DevGuard = KfDeviceGuard;

}  // namespace

struct kf_batch_cache {
    std::vector<std::pair<std::string, BatchShape>> tuned;
    void *table = nullptr;
    size_t table_cap = 0;  // bytes
    DeformState *states = nullptr;
    DeformState *states_host = nullptr;
    int states_cap = 0;
    BlockPartial *parts = nullptr;
    size_t parts_cap = 0;  // slots per parity set
    std::string graph_key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void drop_graph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_key.clear();
    }
};

void kf_batch_release(kf_context *ctx)
{
    kf_batch_cache *b = ctx->batch;
    if (!b)
        return;
    b->drop_graph();
    if (b->table)
        (void)hipFree(b->table);
    if (b->states)
        (void)hipFree(b->states);
    if (b->states_host)
        (void)hipHostFree(b->states_host);
    if (b->parts)
        (void)hipFree(b->parts);
    delete b;
    ctx->batch = nullptr;
}

namespace {

kf_batch_cache *cache_of(kf_context *ctx)
{
    if (!ctx->batch)
        ctx->batch = new kf_batch_cache();
    return ctx->batch;
}

kf_status ensure_workspace(kf_context *ctx, size_t table_bytes, int count, size_t tiles)
{
    kf_batch_cache *b = cache_of(ctx);
    if (table_bytes > b->table_cap) {
        b->drop_graph();
        if (b->table)
            KF_HIP(hipFree(b->table));
        b->table = nullptr;
        b->table_cap = 0;
        KF_HIP(hipMalloc(&b->table, table_bytes));
        b->table_cap = table_bytes;
    }
    if (count > b->states_cap) {
        b->drop_graph();
        if (b->states)
            KF_HIP(hipFree(b->states));
        if (b->states_host)
            KF_HIP(hipHostFree(b->states_host));
        b->states = nullptr;
        b->states_host = nullptr;
        b->states_cap = 0;
        KF_HIP(hipMalloc((void **)&b->states, sizeof(DeformState) * (size_t)count));
        KF_HIP(hipHostMalloc((void **)&b->states_host, sizeof(DeformState) * (size_t)count, hipHostMallocDefault));
        b->states_cap = count;
    }
    if (tiles > b->parts_cap) {
        b->drop_graph();
        if (b->parts)
            KF_HIP(hipFree(b->parts));
        b->parts = nullptr;
        b->parts_cap = 0;
        KF_HIP(hipMalloc((void **)&b->parts, 2 * sizeof(BlockPartial) * tiles));
        b->parts_cap = tiles;
    }
    return KF_OK;
}

std::string shape_key(const kf_volume *const *live, int count)
{
    std::string k = std::to_string((int)live[0]->dtype);
    for (int v = 0; v < count; v++)
        for (int a = 0; a < 3; a++)
            k += ":" + std::to_string(live[v]->dims[a]);
    return k;
}

// Table of the batch for tiles of BLOCK vpt voxels; returns the number of tiles.
template <class S>
size_t build_table(const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u, int count,
                   int vpt, std::vector<BatchVol<S>> &tab)
{
    tab.resize((size_t)count);
    size_t t0 = 0;
    for (int v = 0; v < count; v++) {
        BatchVol<S> &e = tab[(size_t)v];
        e.L = make_read<S>(live[v]);
        e.Cn = make_read<S>(can[v]);
        e.u[0] = u[v]->buf[u[v]->cur];
        e.u[1] = u[v]->buf[u[v]->cur ^ 1];
        e.tile0 = (unsigned int)t0;
        e.ntiles = (unsigned int)(((size_t)live[v]->n + (size_t)BLOCK * vpt - 1) / ((size_t)BLOCK * vpt));
        e.vec4 = live[v]->n % 4 == 0 ? 1 : 0;
        e.pad = 0;
        t0 += e.ntiles;
    }
    return t0;
}

template <class S>
kf_status launch_step(kf_context *ctx, const BatchShape &sh, const BatchVol<S> *dtab, int count, unsigned int total,
                      const AdjConsts &c, DeformState *states, BlockPartial *parts, unsigned int cap, int iter)
{
    const unsigned int grid = sh.blocks ? sh.blocks : total;
    hipStream_t st = ctx->stream;
#define KF_BATCH_LAUNCH(V)                                                                                    \
    hipLaunchKernelGGL((k_adj_batch<S, V>), dim3(grid), dim3(BLOCK), 0, st, dtab, count, total, c, states, parts, \
                       cap, iter)
    switch (sh.vpt) {
    case 1: KF_BATCH_LAUNCH(1); break;
    case 2: KF_BATCH_LAUNCH(2); break;
    case 4: KF_BATCH_LAUNCH(4); break;
    case 8: KF_BATCH_LAUNCH(8); break;
    default: kf_set_error("unsupported batch tile shape"); return KF_E_INVALID_ARG;
    }
#undef KF_BATCH_LAUNCH
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status reset_states(kf_context *ctx, int count, double stop)
{
    hipLaunchKernelGGL(k_batch_reset, dim3((count + 63) / 64), dim3(64), 0, ctx->stream, ctx->batch->states, count,
                       stop);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

// Times every candidate shape: launches of iteration 0, which read u[b]'s value and write
// only its scratch buffer.
template <class S>
kf_status tune(kf_context *ctx, const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u,
               int count, const kf_params *p, BatchShape &best)
{
    const AdjConsts c = kf_adj_consts(p);
    std::vector<BatchVol<S>> tab;
    size_t max_tiles = 0;
    for (int vpt : VPTS)
        max_tiles = std::max(max_tiles, build_table<S>(live, can, u, count, vpt, tab));
    kf_status s = ensure_workspace(ctx, sizeof(BatchVol<S>) * (size_t)count, count, max_tiles);
    if (s != KF_OK)
        return s;
    kf_batch_cache *b = ctx->batch;
    hipEvent_t e0, e1;
    KF_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        kf_set_error("hipEventCreate failed");
        return KF_E_HIP;
    }
    constexpr int WARM = 2, REPS = 5;
    best = BatchShape{};
    best.us = 1e30;
    int cands = 0;
    for (int vpt : VPTS) {
        const size_t total = build_table<S>(live, can, u, count, vpt, tab);
        if (s == KF_OK && hipMemcpyAsync(b->table, tab.data(), sizeof(BatchVol<S>) * tab.size(),
                                         hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
            kf_set_error("batch table upload failed");
            s = KF_E_HIP;
        }
        if (s == KF_OK)
            s = reset_states(ctx, count, 0.0);
        unsigned int last_grid = 0;
        for (int mult : GRID_MULT) {
            if (s != KF_OK)
                break;
            BatchShape sh;
            sh.vpt = vpt;
            sh.blocks = mult ? (unsigned int)std::min<size_t>((size_t)mult * ctx->num_cus, total) : 0u;
            const unsigned int grid = sh.blocks ? sh.blocks : (unsigned int)total;
            if (grid == last_grid || (mult && sh.blocks == total))
                continue;
            last_grid = grid;
            const auto *dtab = (const BatchVol<S> *)b->table;
            for (int r = 0; r < WARM && s == KF_OK; r++)
                s = launch_step<S>(ctx, sh, dtab, count, (unsigned int)total, c, b->states, b->parts,
                                   (unsigned int)b->parts_cap, 0);
            if (s == KF_OK && hipEventRecord(e0, ctx->stream) != hipSuccess)
                s = KF_E_HIP;
            for (int r = 0; r < REPS && s == KF_OK; r++)
                s = launch_step<S>(ctx, sh, dtab, count, (unsigned int)total, c, b->states, b->parts,
                                   (unsigned int)b->parts_cap, 0);
            float ms = 0.f;
            if (s == KF_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                               hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
                kf_set_error("timing the batch candidates failed");
                s = KF_E_HIP;
            }
            if (s != KF_OK)
                break;
            cands++;
            sh.us = 1e3 * ms / REPS;
            if (sh.us < best.us)
                best = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    best.candidates = cands;
    if (s == KF_OK)
        KF_HIP(hipStreamSynchronize(ctx->stream));
    return s;
}

template <class S>
kf_status tuned_shape(kf_context *ctx, const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u,
                      int count, const kf_params *p, bool force, BatchShape &sh)
{
    kf_batch_cache *b = cache_of(ctx);
    const std::string key = shape_key(live, count);
    for (auto &e : b->tuned)
        if (e.first == key && !force) {
            sh = e.second;
            return KF_OK;
        }
    kf_status s = tune<S>(ctx, live, can, u, count, p, sh);
    if (s != KF_OK)
        return s;
    for (auto &e : b->tuned)
        if (e.first == key) {
            e.second = sh;
            return KF_OK;
        }
    b->tuned.emplace_back(key, sh);
    return KF_OK;
}

template <class S>
kf_status run_batch(kf_context *ctx, const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u,
                    int count, const kf_params *p, kf_deform_stats *stats)
{
    BatchShape sh;
    kf_status s = tuned_shape<S>(ctx, live, can, u, count, p, false, sh);
    if (s != KF_OK)
        return s;
    std::vector<BatchVol<S>> tab;
    const size_t total = build_table<S>(live, can, u, count, sh.vpt, tab);
    s = ensure_workspace(ctx, sizeof(BatchVol<S>) * (size_t)count, count, total);
    if (s != KF_OK)
        return s;
    kf_batch_cache *b = ctx->batch;
    // The table (with each field's current buffer first) changes from call to call; the
    // graph reads it from the same device address.
    KF_HIP(hipMemcpyAsync(b->table, tab.data(), sizeof(BatchVol<S>) * tab.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    const AdjConsts c = kf_adj_consts(p);
    const int K = p->max_iters;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%d/%d/%u/%d/%zu/%d/%d/%a/%p/%p/%p/%zu", (int)S::dtype, sh.vpt,
                  sh.blocks, count, total, K, ctx->num_cus, p->stop_max_update, b->table, (void *)b->states,
                  (void *)b->parts, b->parts_cap);
    std::string key(buf);
    key.append((const char *)&c, sizeof(c));
    if (key != b->graph_key) {
        b->drop_graph();
        KF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const auto *dtab = (const BatchVol<S> *)b->table;
        s = reset_states(ctx, count, p->stop_max_update);
        for (int k = 0; k < K && s == KF_OK; k++)
            s = launch_step<S>(ctx, sh, dtab, count, (unsigned int)total, c, b->states, b->parts,
                               (unsigned int)b->parts_cap, k);
        if (s == KF_OK && K > 0) {
            hipLaunchKernelGGL(k_batch_finish<S>, dim3(count), dim3(BLOCK), 0, ctx->stream, dtab, b->states,
                               (const BlockPartial *)b->parts, (unsigned int)b->parts_cap, K - 1);
            if (hipGetLastError() != hipSuccess) {
                kf_set_error("batch finish launch failed");
                s = KF_E_HIP;
            }
        }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        if (s != KF_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return s;
        }
        if (ec != hipSuccess || !g) {
            kf_set_error(std::string("batch graph capture: ") + hipGetErrorString(ec));
            return KF_E_HIP;
        }
        b->graph = g;
        const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            b->drop_graph();
            kf_set_error(std::string("batch graph instantiate: ") + hipGetErrorString(ei));
            return KF_E_HIP;
        }
        b->graph_key = key;
    }
    KF_HIP(hipGraphLaunch(b->exec, ctx->stream));
    KF_HIP(hipMemcpyAsync(b->states_host, b->states, sizeof(DeformState) * (size_t)count, hipMemcpyDeviceToHost,
                          ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    kf_status worst = KF_OK;
    std::string msg;
    for (int v = 0; v < count; v++) {
        const DeformState h = b->states_host[v];
        u[v]->cur ^= (int)(h.iters & 1u);
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
            o.diverged_index = h.diverged == LLONG_MAX ? -1 : (int64_t)h.diverged;
            o.max_update = h.last_max;
            o.active_voxels = h.last_active;
            o.sweep = KF_SWEEP_DENSE;
        }
    }
    if (stats) {
        for (int v = 0; v < count; v++) {
            if (stats[v].diverged_index >= 0)
                continue;
            s = kf_energy_forward(live[v], can[v], u[v], p, &stats[v].energy, nullptr);
            if (s != KF_OK)
                return s;
        }
    }
    if (worst != KF_OK)
        kf_set_error(msg);
    return worst;
}

// Shared argument checks; sets *batched when the persistent adjoint path applies.
kf_status check_batch(const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u, int32_t count,
                      const kf_params *p, bool *batched)
{
    KF_CHECK_ARG(live && can && u && count >= 0, "null argument");
    KF_CHECK_ARG(kf_valid_params(p), "invalid params");
    bool ok = p->method == KF_METHOD_ADJOINT && count > 0;
    for (int32_t v = 0; v < count; v++) {
        kf_status s = kf_check_triplet(live[v], can[v], u[v]);
        if (s != KF_OK)
            return s;
        for (int32_t w = 0; w < v; w++)
            KF_CHECK_ARG(u[w] != u[v], "a field appears twice in the batch");
        ok = ok && live[v]->ctx == live[0]->ctx && live[v]->dtype == live[0]->dtype &&
             3 * live[v]->n < ((int64_t)1 << 30);
    }
    *batched = ok;
    return KF_OK;
}

}  // namespace

extern "C" {

kf_status kf_batch_deform(const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u,
                          int32_t count, const kf_params *p, kf_deform_stats *stats)
{
    bool batched = false;
    kf_status s = check_batch(live, can, u, count, p, &batched);
    if (s != KF_OK)
        return s;
    if (!batched) {
        kf_status worst = KF_OK;
        for (int32_t v = 0; v < count; v++) {
            s = kf_deform(live[v], can[v], u[v], p, stats ? &stats[v] : nullptr);
            if (s == KF_E_DIVERGED)
                worst = s;
            else if (s != KF_OK)
                return s;
        }
        return worst;
    }
    kf_context *ctx = live[0]->ctx;
    DevGuard dg(ctx->device);
    if (p->sweep != KF_SWEEP_DENSE && p->max_iters > 0) {
        bool used = false;
        s = kf_batch_active(ctx, live, can, u, count, p, stats, p->sweep == KF_SWEEP_ACTIVE, &used);
        if (s != KF_OK || used)
            return s;
    }
    KF_DISPATCH_DTYPE(live[0]->dtype, S, return run_batch<S>(ctx, live, can, u, count, p, stats));
    return KF_OK;
}

kf_status kf_batch_autotune(const kf_volume *const *live, const kf_volume *const *can, kf_field *const *u,
                            int32_t count, const kf_params *p, int32_t force, kf_batch_config *best)
{
    bool batched = false;
    kf_status s = check_batch(live, can, u, count, p, &batched);
    if (s != KF_OK)
        return s;
    KF_CHECK_ARG(batched, "the batched launch needs the adjoint method, one context and dtype, and 3 N < 2^30");
    kf_context *ctx = live[0]->ctx;
    DevGuard dg(ctx->device);
    BatchShape sh;
    if (force == 2) {
        // Pin the shape given in *best (no timing).
        KF_CHECK_ARG(best && (best->voxels_per_thread == 1 || best->voxels_per_thread == 2 ||
                              best->voxels_per_thread == 4 || best->voxels_per_thread == 8) &&
                         best->blocks >= 0,
                     "force = 2 needs a shape in *best (voxels_per_thread 1, 2, 4 or 8; blocks >= 0)");
        sh.vpt = best->voxels_per_thread;
        sh.blocks = (unsigned int)best->blocks;
        kf_batch_cache *b = cache_of(ctx);
        const std::string key = shape_key(live, count);
        bool found = false;
        for (auto &e : b->tuned)
            if (e.first == key) {
                e.second = sh;
                found = true;
            }
        if (!found)
            b->tuned.emplace_back(key, sh);
    } else {
        KF_DISPATCH_DTYPE(live[0]->dtype, S, s = tuned_shape<S>(ctx, live, can, u, count, p, force != 0, sh));
        if (s != KF_OK)
            return s;
    }
    if (best) {
        std::memset(best, 0, sizeof(*best));
        best->voxels_per_thread = sh.vpt;
        std::vector<BatchVol<StoreF32>> tab;  // tile counts do not depend on the dtype
        const size_t total = build_table<StoreF32>(live, can, u, count, sh.vpt, tab);
        best->blocks = (int32_t)(sh.blocks && sh.blocks < total ? sh.blocks : total);
        best->candidates = sh.candidates;
        best->us_per_iteration = sh.us;
    }
    return KF_OK;
}

}  // extern "C"
