// k_mc.hip -- marching cubes with the reference's output order.
//
// SDF::getMesh (SDF.cpp:396-414) visits every cell (x, y, z) in z -> y -> x order, skips it
// when its corner (x, y, z) is > MaxSurfaceVoxelDistance or < -UnknownClipDistance, and
// appends each cell's triangles (Polygonise, MarchingCubes.h:366-422) as three fresh
// vertices with y negated (SimpleMesh.h:66-80).  The GPU keeps that order with an ordered
// compaction instead of atomics:
//   1. count: each block owns CHUNK consecutive cells and sums their triangle counts;
//   2. scan: one block turns the per-block totals into exclusive offsets;
//   3. emit: each block recounts its cells round by round (BLOCK cells per round, in cell
//      order), prefix-sums the counts across the block and writes every triangle at its
//      final position.
// Corner values are compared and interpolated in FP64 like the reference (built with
// -ffp-contract=off), so counts, order and vertex values match an FP64 evaluation of the
// same stored grid.
#include "kf_internal.hpp"
#include "kf_mc_cases.h"

#include <vector>

namespace {

constexpr int BLOCK = 256;
constexpr int ITEMS = 16;
constexpr int CHUNK = BLOCK * ITEMS;

__constant__ unsigned char c_tri_count[256];
__constant__ unsigned char c_tri_edges[256][15];
__constant__ unsigned short c_edge_mask[256];

// Corner k of cell (x, y, z): ProcessVolumeCell, SDF.cpp:338-355.
__constant__ signed char c_corner[8][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                           {1, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
__constant__ unsigned char c_edge_corners[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// Inclusive scan of v over the block in thread order, and the block total: a wavefront scan,
// then the wavefront totals through LDS (`s`, BLOCK / 64 entries; two barriers).
__device__ inline unsigned int block_scan_incl(unsigned int v, unsigned int *s, unsigned int &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int t = __shfl_up(v, o, 64);
        if (lane >= o)
            v += t;
    }
    if (lane == 63)
        s[w] = v;
    __syncthreads();
    unsigned int wbase = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < BLOCK / 64; q++) {
        wbase += q < w ? s[q] : 0u;
        tot += s[q];
    }
    __syncthreads();
    total = tot;
    return wbase + v;
}

template <class S>
struct Cells {
    const typename S::T *phi;
    GridDims g;
    float scale;
    double oob, max_surface, clip, iso;
    __device__ inline double at(int x, int y, int z) const
    {
        return g.in(x, y, z) ? (double)S::load(phi, g.idx(x, y, z), scale) : oob;
    }
    // Returns the cube index, or -1 when the cell is skipped or has no triangles.
    __device__ inline int classify(int64_t i, double val[8], int &x, int &y, int &z) const
    {
        kf_xyz(g, i, x, y, z);
        const double d = at(x, y, z);
        if (d > max_surface || d < -clip)
            return -1;
        int ci = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            val[k] = at(x + c_corner[k][0], y + c_corner[k][1], z + c_corner[k][2]);
            if (val[k] < iso)
                ci |= 1 << k;
        }
        return c_edge_mask[ci] ? ci : -1;
    }
};

template <class S>
__global__ __launch_bounds__(BLOCK) void k_mc_count(Cells<S> C, unsigned long long *block_tot)
{
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    unsigned int cnt = 0;
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = base + (int64_t)k * BLOCK + threadIdx.x;
        if (i < C.g.n) {
            double val[8];
            int x, y, z;
            const int ci = C.classify(i, val, x, y, z);
            if (ci >= 0)
                cnt += c_tri_count[ci];
        }
    }
    __shared__ unsigned int s[BLOCK];
    s[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        block_tot[blockIdx.x] = s[0];
}

// Exclusive scan of nb block totals in one block of 1024 threads; offs[nb] = total.
__global__ __launch_bounds__(1024) void k_mc_scan(const unsigned long long *tot, unsigned long long *offs, int nb)
{
    __shared__ unsigned long long s[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    unsigned long long sum = 0;
    for (int b = b0; b < b1; b++)
        sum += tot[b];
    s[t] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan.
    for (int o = 1; o < 1024; o <<= 1) {
        unsigned long long v = (t >= o) ? s[t - o] : 0ull;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    unsigned long long run = s[t] - sum;
    for (int b = b0; b < b1; b++) {
        offs[b] = run;
        run += tot[b];
    }
    if (t == 1023)
        offs[nb] = s[1023];
}

template <class S>
__global__ __launch_bounds__(BLOCK) void k_mc_emit(Cells<S> C, const unsigned long long *offs, float *verts,
                                                   unsigned long long cap)
{
    __shared__ unsigned int s[BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    unsigned long long run = offs[blockIdx.x];
    const double hx = (double)C.g.nx / 2, hy = (double)C.g.ny / 2, hz = (double)C.g.nz / 2;
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = base + (int64_t)k * BLOCK + threadIdx.x;
        double val[8];
        int x = 0, y = 0, z = 0, ci = -1;
        unsigned int cnt = 0;
        if (i < C.g.n) {
            ci = C.classify(i, val, x, y, z);
            if (ci >= 0)
                cnt = c_tri_count[ci];
        }
        // Block-wide scan of cnt in thread order (= cell order within the round).
        unsigned int round_total;
        const unsigned int incl = block_scan_incl(cnt, s, round_total);
        if (cnt) {
            // VertexInterp (MarchingCubes.h:326-356) on the edges this case cuts.
            double p[8][3];
#pragma unroll
            for (int c = 0; c < 8; c++) {
                p[c][0] = (double)(x + c_corner[c][0]) / hx - 1;
                p[c][1] = (double)(y + c_corner[c][1]) / hy - 1;
                p[c][2] = (double)(z + c_corner[c][2]) / hz - 1;
            }
            double vl[12][3];
            const unsigned int em = c_edge_mask[ci];
#pragma unroll
            for (int e = 0; e < 12; e++) {
                if (em & (1u << e)) {
                    const int a = c_edge_corners[e][0], b = c_edge_corners[e][1];
                    const double sc = (C.iso - val[a]) / (val[a] - val[b]);
                    vl[e][0] = p[a][0] + sc * (p[a][0] - p[b][0]);
                    vl[e][1] = p[a][1] + sc * (p[a][1] - p[b][1]);
                    vl[e][2] = p[a][2] + sc * (p[a][2] - p[b][2]);
                }
            }
            unsigned long long t0 = run + (incl - cnt);
            for (unsigned int t = 0; t < cnt; t++) {
                const unsigned long long tri = t0 + t;
                if (tri >= cap)
                    break;
                float *o = verts + tri * 9;
                for (int v = 0; v < 3; v++) {
                    const int e = c_tri_edges[ci][t * 3 + v];
                    o[v * 3 + 0] = (float)vl[e][0];
                    o[v * 3 + 1] = (float)(-vl[e][1]);
                    o[v * 3 + 2] = (float)vl[e][2];
                }
            }
        }
        run += round_total;
    }
}

// ---------------------------------------------------------------------------
// Indexed (welded) output: one vertex per cut grid edge that a visited cell uses.
// Grid edge (node (x, y, z), axis a) has id 3 ((z (ny + 1) + y) (nx + 1) + x) + a over the
// (nx + 1)(ny + 1)(nz + 1) nodes cells can touch; vertices are numbered in id order.
// ---------------------------------------------------------------------------

// Local edge e of a cell: its lower node's offset and its axis (from c_corner / c_edge_corners).
__constant__ signed char c_edge_node[12][4];

struct EdgeGrid {
    int64_t nx1, ny1, nz1;  // node counts per axis
    __device__ inline int64_t id(int x, int y, int z, int a) const { return 3 * ((z * ny1 + y) * nx1 + x) + a; }
    __device__ __host__ inline int64_t count() const { return 3 * nx1 * ny1 * nz1; }
};

template <class S>
__global__ __launch_bounds__(BLOCK) void k_mc_edge_flags(Cells<S> C, EdgeGrid E, unsigned char *flags)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= C.g.n)
        return;
    double val[8];
    int x, y, z;
    const int ci = C.classify(i, val, x, y, z);
    if (ci < 0)
        return;
    const unsigned int em = c_edge_mask[ci];
    for (int e = 0; e < 12; e++)
        if (em & (1u << e))
            flags[E.id(x + c_edge_node[e][0], y + c_edge_node[e][1], z + c_edge_node[e][2], c_edge_node[e][3])] = 1;
}

// Per-block totals of CHUNK consecutive flags.
__global__ __launch_bounds__(BLOCK) void k_flag_sums(const unsigned char *flags, int64_t n, unsigned long long *tot)
{
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    unsigned int cnt = 0;
    for (int k = 0; k < ITEMS; k++) {
        const int64_t j = base + (int64_t)k * BLOCK + threadIdx.x;
        if (j < n)
            cnt += flags[j];
    }
    __shared__ unsigned int s[BLOCK];
    s[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        tot[blockIdx.x] = s[0];
}

// ids[j] = number of flags before j (valid where flags[j] is set), in index order.
__global__ __launch_bounds__(BLOCK) void k_flag_ids(const unsigned char *flags, int64_t n,
                                                    const unsigned long long *offs, unsigned int *ids)
{
    __shared__ unsigned int s[BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    unsigned long long run = offs[blockIdx.x];
    for (int k = 0; k < ITEMS; k++) {
        const int64_t j = base + (int64_t)k * BLOCK + threadIdx.x;
        const unsigned int f = j < n ? flags[j] : 0u;
        unsigned int round_total;
        const unsigned int incl = block_scan_incl(f, s, round_total);
        if (f)
            ids[j] = (unsigned int)(run + incl - 1);
        run += round_total;
    }
}

// Vertex of every flagged edge: VertexInterp from the lower node to the upper one
// (MarchingCubes.h:326-356, in the soup's normalised coordinates, y negated) and the unit
// normal of the SDF gradient (central differences at both nodes, interpolated the same way).
template <class S>
__global__ __launch_bounds__(BLOCK) void k_mc_vertices(Cells<S> C, EdgeGrid E, const unsigned char *flags,
                                                       const unsigned int *ids, float *verts, float *normals)
{
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= E.count() || !flags[j])
        return;
    const int a = (int)(j % 3);
    const int64_t node = j / 3;
    const int x = (int)(node % E.nx1), y = (int)((node / E.nx1) % E.ny1), z = (int)(node / (E.nx1 * E.ny1));
    const int q[3] = {x + (a == 0), y + (a == 1), z + (a == 2)};
    const double v0 = C.at(x, y, z), v1 = C.at(q[0], q[1], q[2]);
    const double hx = (double)C.g.nx / 2, hy = (double)C.g.ny / 2, hz = (double)C.g.nz / 2;
    const double p0[3] = {(double)x / hx - 1, (double)y / hy - 1, (double)z / hz - 1};
    const double p1[3] = {(double)q[0] / hx - 1, (double)q[1] / hy - 1, (double)q[2] / hz - 1};
    const double sc = (C.iso - v0) / (v0 - v1);
    const unsigned int id = ids[j];
    verts[3 * (size_t)id + 0] = (float)(p0[0] + sc * (p0[0] - p1[0]));
    verts[3 * (size_t)id + 1] = (float)(-(p0[1] + sc * (p0[1] - p1[1])));
    verts[3 * (size_t)id + 2] = (float)(p0[2] + sc * (p0[2] - p1[2]));
    if (normals) {
        const double t = -sc;  // position of the vertex along the edge, 0 at node 0
        double g[3];
        const int n0[3] = {x, y, z};
        for (int b = 0; b < 3; b++) {
            int lo0[3] = {n0[0], n0[1], n0[2]}, hi0[3] = {n0[0], n0[1], n0[2]};
            int lo1[3] = {q[0], q[1], q[2]}, hi1[3] = {q[0], q[1], q[2]};
            lo0[b] -= 1;
            hi0[b] += 1;
            lo1[b] -= 1;
            hi1[b] += 1;
            const double d0 = 0.5 * (C.at(hi0[0], hi0[1], hi0[2]) - C.at(lo0[0], lo0[1], lo0[2]));
            const double d1 = 0.5 * (C.at(hi1[0], hi1[1], hi1[2]) - C.at(lo1[0], lo1[1], lo1[2]));
            g[b] = d0 + t * (d1 - d0);
        }
        // d phi / d p = d phi / d index * h per axis; y negated like the positions
        double nrm[3] = {g[0] * hx, -g[1] * hy, g[2] * hz};
        const double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        const double inv = len > 0 ? 1.0 / len : 0.0;
        for (int b = 0; b < 3; b++)
            normals[3 * (size_t)id + b] = (float)(nrm[b] * inv);
    }
}

// Triangles in the soup's order (same offsets as k_mc_emit), as vertex indices.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_mc_emit_indexed(Cells<S> C, EdgeGrid E, const unsigned long long *offs,
                                                           const unsigned int *ids, unsigned int *tris)
{
    __shared__ unsigned int s[BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * CHUNK;
    unsigned long long run = offs[blockIdx.x];
    for (int k = 0; k < ITEMS; k++) {
        const int64_t i = base + (int64_t)k * BLOCK + threadIdx.x;
        double val[8];
        int x = 0, y = 0, z = 0, ci = -1;
        unsigned int cnt = 0;
        if (i < C.g.n) {
            ci = C.classify(i, val, x, y, z);
            if (ci >= 0)
                cnt = c_tri_count[ci];
        }
        unsigned int round_total;
        const unsigned int incl = block_scan_incl(cnt, s, round_total);
        const unsigned long long t0 = run + (incl - cnt);
        for (unsigned int t = 0; t < cnt; t++)
            for (int v = 0; v < 3; v++) {
                const int e = c_tri_edges[ci][t * 3 + v];
                tris[3 * (t0 + t) + v] =
                    ids[E.id(x + c_edge_node[e][0], y + c_edge_node[e][1], z + c_edge_node[e][2], c_edge_node[e][3])];
            }
        run += round_total;
    }
}

template <class S>
Cells<S> make_cells(const kf_volume *v, double iso)
{
    Cells<S> C;
    C.phi = (const typename S::T *)v->phi;
    C.g = grid_of(v);
    C.scale = v->scale();
    C.oob = v->oob();
    C.max_surface = v->max_surface;
    C.clip = v->clip;
    C.iso = iso;
    return C;
}

}  // namespace

// Uploads the case table to the current device's constant memory.
kf_status kf_mc_init_tables()
{
    unsigned char cnt[256];
    unsigned char edges[256][15];
    unsigned short em[256];
    for (int ci = 0; ci < 256; ci++) {
        const char *cs = KF_MC_CASES[ci];
        int k = 0;
        for (; cs[k]; k++) {
            const char ch = cs[k];
            edges[ci][k] = (unsigned char)(ch <= '9' ? ch - '0' : ch - 'a' + 10);
        }
        for (int j = k; j < 15; j++)
            edges[ci][j] = 0;
        cnt[ci] = (unsigned char)(k / 3);
        unsigned short m = 0;
        for (int e = 0; e < 12; e++) {
            const int a = (ci >> KF_MC_EDGE_CORNERS[e][0]) & 1, b = (ci >> KF_MC_EDGE_CORNERS[e][1]) & 1;
            if (a != b)
                m |= (unsigned short)(1u << e);
        }
        em[ci] = m;
    }
    KF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tri_count), cnt, sizeof(cnt)));
    KF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_tri_edges), edges, sizeof(edges)));
    KF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_edge_mask), em, sizeof(em)));
    // Lower node and axis of each local edge, from the corner offsets of ProcessVolumeCell.
    static const signed char corner[8][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                             {1, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    signed char en[12][4];
    for (int e = 0; e < 12; e++) {
        const int a = KF_MC_EDGE_CORNERS[e][0], b = KF_MC_EDGE_CORNERS[e][1];
        int axis = 0;
        for (int d = 0; d < 3; d++) {
            en[e][d] = (signed char)(corner[a][d] < corner[b][d] ? corner[a][d] : corner[b][d]);
            if (corner[a][d] != corner[b][d])
                axis = d;
        }
        en[e][3] = (signed char)axis;
    }
    KF_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_edge_node), en, sizeof(en)));
    return KF_OK;
}

kf_status kf_mc_run(const kf_volume *v, double iso, uint64_t *n_tris, float *verts_host, uint64_t cap)
{
    kf_context *ctx = v->ctx;
    const int nb = (int)((v->n + CHUNK - 1) / CHUNK);
    unsigned long long *tot = nullptr, *offs = nullptr;
    KF_HIP(hipMallocAsync((void **)&tot, sizeof(unsigned long long) * (size_t)(2 * nb + 1), ctx->stream));
    offs = tot + nb;
    kf_status st = KF_OK;
    unsigned long long total = 0;
    KF_DISPATCH_DTYPE(v->dtype, S, {
        Cells<S> C = make_cells<S>(v, iso);
        hipLaunchKernelGGL((k_mc_count<S>), dim3(nb), dim3(BLOCK), 0, ctx->stream, C, tot);
        hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, ctx->stream, (const unsigned long long *)tot, offs, nb);
        hipError_t e = hipMemcpyAsync(&total, offs + nb, sizeof(total), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            kf_set_error(std::string("marching cubes count: ") + hipGetErrorString(e));
            st = KF_E_HIP;
        }
        if (st == KF_OK) {
            *n_tris = total;
            if (verts_host && total > 0) {
                if (cap < total) {
                    kf_set_error("marching cubes: capacity below triangle count");
                    st = KF_E_CAPACITY;
                } else {
                    float *dverts = nullptr;
                    e = hipMallocAsync((void **)&dverts, sizeof(float) * 9 * (size_t)total, ctx->stream);
                    if (e == hipSuccess) {
                        hipLaunchKernelGGL((k_mc_emit<S>), dim3(nb), dim3(BLOCK), 0, ctx->stream, C,
                                           (const unsigned long long *)offs, dverts, (unsigned long long)total);
                        e = hipGetLastError();
                        if (e == hipSuccess)
                            e = hipMemcpyAsync(verts_host, dverts, sizeof(float) * 9 * (size_t)total,
                                               hipMemcpyDeviceToHost, ctx->stream);
                        (void)hipFreeAsync(dverts, ctx->stream);
                        if (e == hipSuccess)
                            e = hipStreamSynchronize(ctx->stream);
                    }
                    if (e != hipSuccess) {
                        kf_set_error(std::string("marching cubes emit: ") + hipGetErrorString(e));
                        st = e == hipErrorOutOfMemory ? KF_E_OOM : KF_E_HIP;
                    }
                }
            }
        }
    });
    (void)hipFreeAsync(tot, ctx->stream);
    return st;
}

kf_status kf_mc_run_indexed(const kf_volume *v, double iso, uint64_t *n_verts, uint64_t *n_tris, float *verts_host,
                            float *normals_host, uint32_t *tris_host, uint64_t cap_verts, uint64_t cap_tris)
{
    kf_context *ctx = v->ctx;
    hipStream_t st = ctx->stream;
    const int nb = (int)((v->n + CHUNK - 1) / CHUNK);
    EdgeGrid E{(int64_t)v->dims[0] + 1, (int64_t)v->dims[1] + 1, (int64_t)v->dims[2] + 1};
    const int64_t ne = E.count();
    const int nbe = (int)((ne + CHUNK - 1) / CHUNK);
    // One scratch block: triangle block totals/offsets, edge block totals/offsets, edge flags.
    const size_t tri_words = (size_t)(2 * nb + 1), edge_words = (size_t)(2 * nbe + 1);
    char *scratch = nullptr;
    KF_HIP(hipMallocAsync((void **)&scratch, 8 * (tri_words + edge_words) + (size_t)ne, st));
    unsigned long long *ttot = (unsigned long long *)scratch, *toffs = ttot + nb;
    unsigned long long *etot = ttot + tri_words, *eoffs = etot + nbe;
    unsigned char *flags = (unsigned char *)(etot + edge_words);
    kf_status status = KF_OK;
    unsigned long long counts[2] = {0, 0};
    hipError_t e = hipSuccess;
    unsigned int *ids = nullptr, *dtris = nullptr;
    float *dverts = nullptr, *dnorm = nullptr;
    KF_DISPATCH_DTYPE(v->dtype, S, {
        Cells<S> C = make_cells<S>(v, iso);
        hipLaunchKernelGGL((k_mc_count<S>), dim3(nb), dim3(BLOCK), 0, st, C, ttot);
        hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, st, (const unsigned long long *)ttot, toffs, nb);
        e = hipMemsetAsync(flags, 0, (size_t)ne, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL((k_mc_edge_flags<S>), dim3((unsigned int)((v->n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                               st, C, E, flags);
            hipLaunchKernelGGL(k_flag_sums, dim3(nbe), dim3(BLOCK), 0, st, (const unsigned char *)flags, ne, etot);
            hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, st, (const unsigned long long *)etot, eoffs, nbe);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(&counts[0], toffs + nb, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&counts[1], eoffs + nbe, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            kf_set_error(std::string("marching cubes (indexed) count: ") + hipGetErrorString(e));
            status = KF_E_HIP;
        }
        const unsigned long long ntri = counts[0], nvert = counts[1];
        if (status == KF_OK) {
            *n_tris = ntri;
            *n_verts = nvert;
            const bool want = (verts_host || normals_host || tris_host) && ntri > 0;
            if (want && ((verts_host || normals_host) && cap_verts < nvert)) {
                kf_set_error("marching cubes (indexed): vertex capacity below the vertex count");
                status = KF_E_CAPACITY;
            } else if (want && tris_host && cap_tris < ntri) {
                kf_set_error("marching cubes (indexed): triangle capacity below the triangle count");
                status = KF_E_CAPACITY;
            } else if (want) {
                e = hipMallocAsync((void **)&ids, sizeof(unsigned int) * (size_t)ne, st);
                if (e == hipSuccess)
                    e = hipMallocAsync((void **)&dverts, sizeof(float) * 3 * (size_t)nvert, st);
                if (e == hipSuccess && normals_host)
                    e = hipMallocAsync((void **)&dnorm, sizeof(float) * 3 * (size_t)nvert, st);
                if (e == hipSuccess)
                    e = hipMallocAsync((void **)&dtris, sizeof(unsigned int) * 3 * (size_t)ntri, st);
                if (e == hipSuccess) {
                    hipLaunchKernelGGL(k_flag_ids, dim3(nbe), dim3(BLOCK), 0, st, (const unsigned char *)flags, ne,
                                       (const unsigned long long *)eoffs, ids);
                    hipLaunchKernelGGL((k_mc_vertices<S>), dim3((unsigned int)((ne + BLOCK - 1) / BLOCK)), dim3(BLOCK),
                                       0, st, C, E, (const unsigned char *)flags, (const unsigned int *)ids, dverts,
                                       dnorm);
                    hipLaunchKernelGGL((k_mc_emit_indexed<S>), dim3(nb), dim3(BLOCK), 0, st, C, E,
                                       (const unsigned long long *)toffs, (const unsigned int *)ids, dtris);
                    e = hipGetLastError();
                }
                if (e == hipSuccess && verts_host)
                    e = hipMemcpyAsync(verts_host, dverts, sizeof(float) * 3 * (size_t)nvert, hipMemcpyDeviceToHost, st);
                if (e == hipSuccess && normals_host)
                    e = hipMemcpyAsync(normals_host, dnorm, sizeof(float) * 3 * (size_t)nvert, hipMemcpyDeviceToHost,
                                       st);
                if (e == hipSuccess && tris_host)
                    e = hipMemcpyAsync(tris_host, dtris, sizeof(unsigned int) * 3 * (size_t)ntri, hipMemcpyDeviceToHost,
                                       st);
                if (e == hipSuccess)
                    e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                    kf_set_error(std::string("marching cubes (indexed) emit: ") + hipGetErrorString(e));
                    status = e == hipErrorOutOfMemory ? KF_E_OOM : KF_E_HIP;
                }
            }
        }
    });
    for (void *ptr : {(void *)ids, (void *)dverts, (void *)dnorm, (void *)dtris})
        if (ptr)
            (void)hipFreeAsync(ptr, st);
    (void)hipFreeAsync(scratch, st);
    return status;
}
