// k_mc.hip -- marching cubes with the reference's output order.
//
// SDF::getMesh (SDF.cpp:396-414) visits every cell (x, y, z) in z -> y -> x order, skips it
// when its corner (x, y, z) is > MaxSurfaceVoxelDistance or < -UnknownClipDistance, and
// appends each cell's triangles (Polygonise, MarchingCubes.h:366-422) as three fresh
// vertices with y negated (SimpleMesh.h:66-80).  The GPU keeps that order with an ordered
// compaction instead of atomics:
//   1. count: a thread classifies CPT consecutive cells (the node values it needs come in as
//      vector loads, all issued before any is used); a wavefront owns UNIT consecutive cells,
//      a workgroup CELLS, and each writes its triangle count;
//   2. scan: one block turns the workgroup counts into exclusive offsets and the total, and
//      lists the workgroups whose cells hold triangles;
//   3. emit: a workgroup per unit of a listed block classifies the unit's cells again,
//      prefix-sums the thread counts and writes every triangle at its final position; only
//      the cells that hold triangles interpolate.
// The counts and offsets of the last volume counted are kept (kf_mc_cache) until the volume
// changes, so the reference-style size query followed by the extraction counts once.
// Corner values are compared and interpolated in FP64 like the reference (built with
// -ffp-contract=off), so counts, order and vertex values match an FP64 evaluation of the
// same stored grid.
#include "kf_internal.hpp"
#include "kf_mc_cases.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int BLOCK = 256;
constexpr int CPT = 8;               // cells per thread (a multiple of 4)
constexpr int CELLS = BLOCK * CPT;   // cells per workgroup of the count pass
constexpr int UNIT = 64 * CPT;       // cells per wavefront of the count pass: one unit
constexpr int UNITS_PER_BLOCK = CELLS / UNIT;
static_assert(CELLS == KF_MC_CELLS && UNIT == KF_MC_UNIT, "cells per count workgroup and per unit");
constexpr int ITEMS = 16;
constexpr int FLAG_CHUNK = BLOCK * ITEMS;  // edge flags per block of the indexed variant's scan

__constant__ unsigned char c_tri_count[256];
__constant__ __align__(16) unsigned char c_tri_edges[256][15];
__constant__ unsigned short c_edge_mask[256];

// Corner k of cell (x, y, z): ProcessVolumeCell, SDF.cpp:338-355 (compile-time copies for
// the unrolled loops of the count and emit passes).
constexpr int kCornerX[8] = {1, 0, 0, 1, 1, 0, 0, 1};
constexpr int kCornerY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
constexpr int kCornerZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__constant__ signed char c_corner[8][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                           {1, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
__constant__ unsigned char c_edge_corners[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// Inclusive scan of v over the block (NT threads) in thread order, and the block total: a
// wavefront scan, then the wavefront totals through LDS (`s`, NT / 64 entries; two barriers).
template <int NT>
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
    for (int q = 0; q < NT / 64; q++) {
        wbase += q < w ? s[q] : 0u;
        tot += s[q];
    }
    __syncthreads();
    total = tot;
    return wbase + v;
}

// Stored SDF values as raw floats: the value before the decode scale (F32 the value itself,
// BF16 its bits as the high half of a float, F16 converted), so that decode = raw * scale for
// BF16 and F16 and decode = raw for F32.
template <class S>
__device__ inline float raw_lo(unsigned int w)
{
    if constexpr (S::dtype == KF_BF16)
        return __uint_as_float(w << 16);
    else
        return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
}
template <class S>
__device__ inline float raw_hi(unsigned int w)
{
    if constexpr (S::dtype == KF_BF16)
        return __uint_as_float(w & 0xffff0000u);
    else
        return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16));
}
template <class S>
__device__ inline float raw1(const typename S::T *p, int64_t i)
{
    if constexpr (S::dtype == KF_F32)
        return p[i];
    else if constexpr (S::dtype == KF_BF16)
        return __uint_as_float((unsigned int)p[i] << 16);
    else
        return (float)p[i];
}
// Eight consecutive raw values from a 16-byte aligned p: one 16-byte load (two for F32).
template <class S>
__device__ inline void raw8(const typename S::T *p, float o[8])
{
    if constexpr (S::dtype == KF_F32) {
        const float4 a = reinterpret_cast<const float4 *>(p)[0], b = reinterpret_cast<const float4 *>(p)[1];
        o[0] = a.x;
        o[1] = a.y;
        o[2] = a.z;
        o[3] = a.w;
        o[4] = b.x;
        o[5] = b.y;
        o[6] = b.z;
        o[7] = b.w;
    } else {
        const uint4 q = *reinterpret_cast<const uint4 *>(p);
        const unsigned int w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o[2 * k] = raw_lo<S>(w[k]);
            o[2 * k + 1] = raw_hi<S>(w[k]);
        }
    }
}

// Corner k of a cell in the node-major bit order of Cells::count8: bit r + 4 dx for the node
// at x + dx in node row r = dy + 2 dz (corners as in kCornerX/Y/Z).
constexpr int kNodeBit[8] = {4, 0, 1, 5, 6, 2, 3, 7};

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
    // Cube index of the cell at (x, y, z), or -1 when the cell is skipped or has no triangles;
    // val receives its corner values.
    __device__ inline int classify_at(int x, int y, int z, double val[8]) const
    {
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
        return ci == 0 || ci == 255 ? -1 : ci;  // every other case cuts some edge
    }
    __device__ inline int classify(int64_t i, double val[8], int &x, int &y, int &z) const
    {
        kf_xyz(g, i, x, y, z);
        return classify_at(x, y, z, val);
    }
    // The value at node (x, y, z), or oob outside the grid, with the load issued whatever the
    // node (from node 0 when outside): a caller's loads are then all in flight together
    // instead of one round trip per bounds test.
    __device__ inline double at_flat(int x, int y, int z) const
    {
        const bool inside = g.in(x, y, z);
        const double v = (double)S::load(phi, inside ? g.idx(x, y, z) : 0, scale);
        return inside ? v : oob;
    }
    // The classification compares, done in FP32 with the same outcome as the reference's FP64
    // ones: a stored value decodes to a float f, and for a double X, (double) f < X exactly
    // when f < the least float >= X (iso_up, clip_up), and (double) f > X exactly when f > the
    // greatest float <= X (max_dn); a node outside the grid reads oob, and stands in the FP32
    // test as oob_f (-inf when oob < iso, else +inf), which compares with iso_up as oob does
    // with iso.  NaN fails every comparison either way.
    float iso_up, max_dn, clip_up, oob_f;
    // Cube index from the corner-below bits m and the cell's own node value d (corner 1).
    __device__ inline int case_of(float d, int m) const
    {
        return (d > max_dn || d < clip_up || m == 0 || m == 255) ? -1 : m;
    }
    // The same tests on raw values (see raw1): decode is monotone in the raw value, so for a
    // raw r, decode(r) < iso_up exactly when r < r_iso (the least raw value that decodes to
    // >= iso_up), decode(r) < clip_up exactly when r < r_clip, and decode(r) > max_dn exactly
    // when r > r_max (the greatest raw value that decodes to <= max_dn); the host finds the
    // three thresholds by bisection over the floats.  NaN fails every comparison either way.
    // oob_bit is the below-iso bit of a node outside the grid (oob_f < iso_up).
    float r_iso, r_clip, r_max;
    unsigned int oob_bit;

    // Count pass of a grid whose rows are a multiple of 8 nodes long: the 8 cells i0 .. i0 + 7
    // (i0 a multiple of 8, i0 < n) lie in one row, and each of their four node rows starts
    // 16-byte aligned.  load8 issues the loads (a row past the grid is read from row 0, and
    // node 8 past the row's end from node 7, their values then replaced);
    // count8 packs the below-iso bits node-major, bit r + 4 j for node j of row r, so that
    // cell c's eight corners are the byte at bit 4 c, and sums the triangle counts that the
    // table sc holds for that byte (kNodeBit order).
    __device__ inline void load8(int64_t i0, float v[4][8], float e[4], unsigned int &rows_out, bool &xl) const
    {
        int x0, y0, z0;
        kf_xyz(g, i0, x0, y0, z0);
        xl = x0 + 8 < g.nx;
        // node row r = dy + 2 dz starts at i0 + dy nx + dz nx ny (i0 is node (x0, y0, z0))
        const bool iny = y0 + 1 < g.ny, inz = z0 + 1 < g.nz;
        const int64_t sy = iny ? g.nx : 0, sz = inz ? (int64_t)g.nx * g.ny : 0;
        rows_out = (iny ? 0u : 0xau) | (inz ? 0u : 0xcu);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t b = i0 + (r & 1 ? sy : 0) + (r & 2 ? sz : 0);
            raw8<S>(phi + b, v[r]);
            e[r] = raw1<S>(phi, xl ? b + 8 : b + 7);
        }
    }
    __device__ inline unsigned int count8(const float v[4][8], const float e[4], unsigned int rows_out, bool xl,
                                          const unsigned char *sc) const
    {
        unsigned int w = 0, w8 = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (int j = 0; j < 8; j++)
                w |= v[r][j] < r_iso ? 1u << (4 * j + r) : 0u;
            w8 |= e[r] < r_iso ? 1u << r : 0u;
        }
        // nodes past the grid read oob: whole rows, and node 8 past the row's end
        const unsigned int mw = rows_out * 0x11111111u, m8 = xl ? rows_out : 0xfu;
        w = oob_bit ? w | mw : w & ~mw;
        w8 = oob_bit ? w8 | m8 : w8 & ~m8;
        unsigned int cnt = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const unsigned int m = c < 7 ? (w >> (4 * c)) & 0xffu : (w >> 28) | (w8 << 4);
            const float d = v[0][c];  // corner 1, node (x0 + c, y0, z0), inside the grid
            const unsigned int t = sc[m];
            cnt += d > r_max || d < r_clip ? 0u : t;
        }
        return cnt;
    }
    // Cube indices (-1: skipped, no triangles or past the grid) and coordinates of the N
    // cells i0 .. i0 + N - 1.  When they lie in one row, the 4 x (N + 1) node values come from
    // the four node rows (y, z), (y + 1, z), (y, z + 1), (y + 1, z + 1) (vector loads of four,
    // and one scalar each); otherwise (a group that wraps into the next row or runs past the
    // grid) cell by cell.  Either way every load is issued before any is used.  With VALS, also
    // each cell's eight corner values (as decoded; meaningless for a corner past the grid) in
    // cv and, in om, bit k set for a corner k past the grid.
    template <int N, bool VALS = false>
    __device__ inline void quad(int64_t i0, int ci[N], int cx[N], int cy[N], int cz[N], float (*cv)[8] = nullptr,
                                unsigned int *om = nullptr) const
    {
        int x0, y0, z0;
        kf_xyz(g, i0, x0, y0, z0);
        if (i0 + N <= g.n && x0 + N <= g.nx) {
            float v[4][N + 1];
            bool in[4];
            const bool xl = x0 + N < g.nx;  // the x + 1 corner of the last cell is inside
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // a row past the grid is read from the last row instead (and its values not
                // used): every row's loads are then issued unconditionally, together
                const int y = y0 + (r & 1), z = z0 + (r >> 1);
                in[r] = y < g.ny && z < g.nz;
                const int64_t b = g.idx(x0, y < g.ny ? y : g.ny - 1, z < g.nz ? z : g.nz - 1);
                if constexpr (N % 4 == 0) {
#pragma unroll
                    for (int q = 0; q < N; q += 4)
                        kf_load4_unaligned<S>(phi + b + q, scale, &v[r][q]);
                } else {
#pragma unroll
                    for (int q = 0; q < N; q++)
                        v[r][q] = S::load(phi, b + q, scale);
                }
                v[r][N] = S::load(phi, xl ? b + N : b, scale);
            }
#pragma unroll
            for (int c = 0; c < N; c++) {
                cx[c] = x0 + c;
                cy[c] = y0;
                cz[c] = z0;
                // corner k = (x + dx, y + dy, z + dz) is node row dy + 2 dz, column c + dx
                int m = 0;
                unsigned int out = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int r = kCornerY[k] + 2 * kCornerZ[k];
                    const bool inside = in[r] && (c + kCornerX[k] < N || xl);
                    const float f = inside ? v[r][c + kCornerX[k]] : oob_f;
                    m |= f < iso_up ? 1 << k : 0;
                    if constexpr (VALS) {
                        cv[c][k] = v[r][c + kCornerX[k]];
                        out |= inside ? 0u : 1u << k;
                    }
                }
                if constexpr (VALS)
                    om[c] = out;
                ci[c] = case_of(v[0][c], m);
            }
        } else {
            int x = x0, y = y0, z = z0;
#pragma unroll
            for (int c = 0; c < N; c++) {
                if (x == g.nx) {
                    x = 0;
                    if (++y == g.ny) {
                        y = 0;
                        ++z;
                    }
                }
                cx[c] = x;
                cy[c] = y;
                cz[c] = z;
                int m = 0;
                unsigned int out = 0;
                float d = 0.0f;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int xk = x + kCornerX[k], yk = y + kCornerY[k], zk = z + kCornerZ[k];
                    const bool inside = g.in(xk, yk, zk);
                    const float f = S::load(phi, inside ? g.idx(xk, yk, zk) : 0, scale);
                    if (k == 1)
                        d = f;  // corner 1 is (x, y, z), inside for every cell of the grid
                    m |= (inside ? f : oob_f) < iso_up ? 1 << k : 0;
                    if constexpr (VALS) {
                        cv[c][k] = f;
                        out |= inside ? 0u : 1u << k;
                    }
                }
                if constexpr (VALS)
                    om[c] = out;
                ci[c] = i0 + c >= g.n ? -1 : case_of(d, m);
                ++x;
            }
        }
    }
};

// Triangles per unit of UNIT cells (a wavefront's: a wavefront sum) and per workgroup of
// CELLS cells (its UNITS_PER_BLOCK unit totals summed through LDS).
static_assert(BLOCK == 256, "one thread per case loads the triangle-count table");

// ROWS8: the grid's rows are a multiple of 8 nodes long (Cells::load8 / count8); otherwise
// the general classification (Cells::quad).
template <class S, bool ROWS8>
__global__ __launch_bounds__(BLOCK) void k_mc_count(Cells<S> C, unsigned int *__restrict__ unit_tot,
                                                    unsigned int *__restrict__ block_tot)
{
    static_assert(CPT == 8, "one load8 group of cells per thread");
    __shared__ unsigned char s_cnt[256];
    __shared__ unsigned int s_w[UNITS_PER_BLOCK];
    const int64_t i0 = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * CPT;
    unsigned int cnt = 0;
    if constexpr (ROWS8) {
        // the count table indexed by node-major corner bits
        unsigned int ci = 0;
#pragma unroll
        for (int k = 0; k < 8; k++)
            ci |= ((threadIdx.x >> kNodeBit[k]) & 1u) << k;
        s_cnt[threadIdx.x] = c_tri_count[ci];
        float v[4][8], e[4];
        unsigned int rows_out = 0;
        bool xl = false;
        const bool live = i0 < C.g.n;
        C.load8(live ? i0 : 0, v, e, rows_out, xl);
        __syncthreads();
        cnt = live ? C.count8(v, e, rows_out, xl, s_cnt) : 0u;
    } else {
        s_cnt[threadIdx.x] = c_tri_count[threadIdx.x];
        int ci[CPT], x[CPT], y[CPT], z[CPT];
        C.template quad<CPT>(i0, ci, x, y, z);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < CPT; c++)
            cnt += ci[c] >= 0 ? s_cnt[ci[c]] : 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) {
        unit_tot[(int64_t)blockIdx.x * UNITS_PER_BLOCK + (threadIdx.x >> 6)] = cnt;
        s_w[threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
#pragma unroll
        for (int u = 0; u < UNITS_PER_BLOCK; u++)
            t += s_w[u];
        block_tot[blockIdx.x] = t;
    }
}

// Exclusive scan of nb block totals in one block of 1024 threads (each sums a run of
// consecutive totals; wavefront scans, then the 16 wavefront totals through LDS); offs[nb] =
// total, also written to *total_out when given.  With blist, also the blocks whose total is
// not zero, ascending: blist[0] = their number, blist[1 ..] = their indices.
__global__ __launch_bounds__(1024) void k_mc_scan(const unsigned int *__restrict__ tot, unsigned long long *offs,
                                                  int nb, unsigned long long *total_out, unsigned int *blist)
{
    __shared__ unsigned long long s[16];
    __shared__ unsigned int s_ne[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (nb + 1023) / 1024;
    const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    // runs of 8 totals, loaded together (indices clamped inside the array, the excess ignored);
    // the first run stays in registers for the offsets
    unsigned int q0[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
        q0[j] = tot[min(b0 + j, nb - 1)];
    unsigned long long sum = 0;
    unsigned int ne = 0;  // blocks of the run with triangles
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (b0 + j < b1) {
            sum += q0[j];
            ne += q0[j] != 0u;
        }
    for (int b = b0 + 8; b < b1; b += 8) {
        unsigned int q[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            q[j] = tot[min(b + j, nb - 1)];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + j < b1) {
                sum += q[j];
                ne += q[j] != 0u;
            }
    }
    unsigned long long v = sum;
    unsigned int vn = ne;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(v, o, 64);
        const unsigned int yn = __shfl_up(vn, o, 64);
        if (lane >= o) {
            v += y;
            vn += yn;
        }
    }
    if (lane == 63) {
        s[w] = v;
        s_ne[w] = vn;
    }
    __syncthreads();
    unsigned long long run = v - sum, all = 0;
    unsigned int pos = vn - ne, nall = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        run += q < w ? s[q] : 0ull;
        all += s[q];
        pos += q < w ? s_ne[q] : 0u;
        nall += s_ne[q];
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (b0 + j < b1) {
            offs[b0 + j] = run;
            run += q0[j];
            if (blist && q0[j] != 0u)
                blist[1 + pos++] = (unsigned int)(b0 + j);
        }
    for (int b = b0 + 8; b < b1; b += 8) {
        unsigned int q[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            q[j] = tot[min(b + j, nb - 1)];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + j < b1) {
                offs[b + j] = run;
                run += q[j];
                if (blist && q[j] != 0u)
                    blist[1 + pos++] = (unsigned int)(b + j);
            }
    }
    if (t == 1023) {
        offs[nb] = all;
        if (total_out)
            *total_out = all;
        if (blist)
            blist[0] = nall;
    }
}

// Threads per workgroup of the emit pass: the UNIT cells of a unit, ECPT per thread, so that
// a surface-dense unit still has a thread for nearly every cell that emits.
constexpr int EB = 512;
constexpr int ECPT = UNIT / EB;
static_assert(ECPT * EB == UNIT, "emit workgroups cover the units");

// The triangles of cell (x, y, z), case ci (Polygonise, MarchingCubes.h:411-419), from
// position t0 on, y negated (SimpleMesh.h:66-72); triangles at or past cap are dropped.  Each
// vertex is VertexInterp (MarchingCubes.h:326-356) on its edge, p1 + (iso - v1) / (v1 - v2)
// (p1 - p2) with p1 the edge's first corner, computed where the triangle uses it (the same
// operations as once per edge, so the same values).  The corner values sit in the thread's
// column of the block's LDS array sv as the floats they decode to (exact), with bit k of
// `outside` set for a corner past the grid, which reads oob (an edge picks its corners by
// runtime index, which registers cannot serve without spilling).  The classification pass
// left them in column col; with GATHER, the thread loads them into that column first.
template <class S, bool GATHER>
__device__ __forceinline__ void emit_cell(const Cells<S> &C, int x, int y, int z, int ci, unsigned int outside,
                                          unsigned long long t0, float *__restrict__ verts, unsigned long long cap,
                                          float (*sv)[EB], int col, const unsigned char *s_edges,
                                          const unsigned char *s_cnt)
{
    const int tid = col;
    if constexpr (GATHER) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int xk = x + kCornerX[k], yk = y + kCornerY[k], zk = z + kCornerZ[k];
            const bool inside = C.g.in(xk, yk, zk);
            sv[k][tid] = S::load(C.phi, inside ? C.g.idx(xk, yk, zk) : 0, C.scale);
        }
    }
    // corner positions in the soup's normalised coordinates (SDF.cpp:338-355: v / (dims / 2) -
    // 1), two per axis: corners 0, 3, 4, 7 are at x + 1 (bits of 0x99), 2, 3, 6, 7 at y + 1
    // (0xcc), 4..7 at z + 1
    const double hx = (double)C.g.nx / 2, hy = (double)C.g.ny / 2, hz = (double)C.g.nz / 2;
    const double px[2] = {(double)x / hx - 1, (double)(x + 1) / hx - 1};
    const double py[2] = {(double)y / hy - 1, (double)(y + 1) / hy - 1};
    const double pz[2] = {(double)z / hz - 1, (double)(z + 1) / hz - 1};
    const unsigned int cnt = s_cnt[ci];
    for (unsigned int t = 0; t < cnt; t++) {
        const unsigned long long tri = t0 + t;
        if (tri >= cap)
            break;
        float *o = verts + tri * 9;
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const int e = s_edges[ci * 15 + t * 3 + v];
            // ends of edge e as Polygonise's VertexInterp calls take them (MarchingCubes.h:386-409):
            // 0..3 join corners e, e + 1 (mod 4); 4..7 join e, e + 1 (4..7 mod 4); 8..11 join e - 8, e - 4
            const int a = e < 8 ? e : e - 8;
            const int b = e < 4 ? (e + 1) & 3 : e < 8 ? 4 + ((e + 1) & 3) : e - 4;
            // both LDS reads are issued whatever the corners (a corner past the grid reads its
            // column's placeholder, then takes oob): no branch, one wait
            float fa = sv[a][tid], fb = sv[b][tid];
            asm volatile("" : "+v"(fa), "+v"(fb));
            const double va = (outside >> a) & 1u ? C.oob : (double)fa;
            const double vb = (outside >> b) & 1u ? C.oob : (double)fb;
            const double pa[3] = {(0x99 >> a) & 1 ? px[1] : px[0], (0xcc >> a) & 1 ? py[1] : py[0],
                                  a >> 2 ? pz[1] : pz[0]};
            const double pb[3] = {(0x99 >> b) & 1 ? px[1] : px[0], (0xcc >> b) & 1 ? py[1] : py[0],
                                  b >> 2 ? pz[1] : pz[0]};
            const double sc = (C.iso - va) / (va - vb);
            o[v * 3 + 0] = (float)(pa[0] + sc * (pa[0] - pb[0]));
            o[v * 3 + 1] = (float)(-(pa[1] + sc * (pa[1] - pb[1])));
            o[v * 3 + 2] = (float)(pa[2] + sc * (pa[2] - pb[2]));
        }
    }
}

// Emit: only the units of the count blocks that hold triangles are visited (blist, from the
// scan), in a grid of at most one workgroup per such unit that strides over them, so the many
// blocks of cells without triangles cost no dispatch, and a surface spreads over as many
// workgroups as it has units.  A workgroup takes unit `part` of listed block q / 4,
// classifies its UNIT cells (ECPT per thread; a unit without triangles lists none), lists
// the ones with triangles in LDS in cell order with their triangle offsets (one block scan of
// the packed counts) and corner values, and then spreads the listed cells over its threads,
// one cell per thread, so the FP64 interpolation of a surface-dense stretch of cells runs in
// parallel rather than one after another in one lane.  The unit's first triangle is its
// block's offset plus the totals of the block's units before it (one 16-byte load beside the
// offset's).  The list length and the first block index are loaded together, then the offset,
// the unit totals and the unit's nodes, and the corner values come from the classification's
// own loads: two round trips to memory before the stores.
template <class S>
__global__ __launch_bounds__(EB) void k_mc_emit(Cells<S> C, const unsigned long long *__restrict__ offs,
                                                const uint4 *__restrict__ unit_tot4,
                                                const unsigned int *__restrict__ blist, float *__restrict__ verts,
                                                unsigned long long cap)
{
    __shared__ unsigned int s[EB / 64];
    __shared__ unsigned char s_cnt[256];
    __shared__ unsigned int s_cell[UNIT];   // (outside corners << 24) | (cell within the unit << 8) | case
    __shared__ unsigned short s_tri[UNIT];  // its first triangle within the unit
    __shared__ float s_val[8][EB];          // corner values of listed cells 0 .. EB - 1
    __shared__ unsigned int s_edges[256 * 15 / 4];  // the case table's edge ids
    static_assert(UNITS_PER_BLOCK == 4, "a block's unit totals are one 16-byte load");
    // blist holds nb + 1 entries and the grid has at most 4 nb workgroups
    const unsigned int nne = blist[0], first = blist[1 + blockIdx.x / 4];
    if (blockIdx.x >= 4 * nne)
        return;
    if (threadIdx.x < 256)
        s_cnt[threadIdx.x] = c_tri_count[threadIdx.x];
    for (int k = threadIdx.x; k < 256 * 15 / 4; k += EB)
        s_edges[k] = reinterpret_cast<const unsigned int *>(&c_tri_edges[0][0])[k];
    for (unsigned int q = blockIdx.x; q < 4 * nne; q += gridDim.x) {
        const unsigned int blk = q == blockIdx.x ? first : blist[1 + q / 4], part = q % 4;
        const uint4 ut = unit_tot4[blk];
        const unsigned int before = (part > 0 ? ut.x : 0u) + (part > 1 ? ut.y : 0u) + (part > 2 ? ut.z : 0u);
        const unsigned long long base = offs[blk] + before;
        const int64_t blk0 = ((int64_t)blk * UNITS_PER_BLOCK + part) * UNIT;
        const int64_t i0 = blk0 + (int64_t)threadIdx.x * ECPT;
        int ci[ECPT], x[ECPT], y[ECPT], z[ECPT];
        float cv[ECPT][8];
        unsigned int om[ECPT];
        C.template quad<ECPT, true>(i0, ci, x, y, z, cv, om);
        __syncthreads();  // the tables are in place, the previous unit's cells emitted
        unsigned int ns = 0, nt = 0;
#pragma unroll
        for (int c = 0; c < ECPT; c++)
            if (ci[c] >= 0) {
                ns += 1;
                nt += s_cnt[ci[c]];
            }
        // cells in the high half, triangles (at most 5 UNIT < 2^16) in the low half
        static_assert(5 * UNIT < (1 << 16), "packed counts");
        unsigned int total;
        const unsigned int incl = block_scan_incl<EB>((ns << 16) | nt, s, total);
        unsigned int slot = (incl >> 16) - ns, tri = (incl & 0xffffu) - nt;
#pragma unroll
        for (int c = 0; c < ECPT; c++)
            if (ci[c] >= 0) {
                s_cell[slot] = (om[c] << 24) | ((unsigned int)(threadIdx.x * ECPT + c) << 8) | (unsigned int)ci[c];
                s_tri[slot] = (unsigned short)tri;
                if (slot < EB) {
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        s_val[k][slot] = cv[c][k];
                }
                slot += 1;
                tri += s_cnt[ci[c]];
            }
        __syncthreads();
        const unsigned int ncells = total >> 16;
        // listed cell e: its values in column e for e < EB; past that (a unit with more
        // listed cells than threads) the thread gathers them into its own column, which its
        // first cell no longer needs
        const unsigned char *edges = reinterpret_cast<const unsigned char *>(s_edges);
        if (threadIdx.x < ncells) {
            const unsigned int e = threadIdx.x, cell = s_cell[e];
            int cx, cy, cz;
            kf_xyz(C.g, blk0 + ((cell >> 8) & 0xffffu), cx, cy, cz);
            emit_cell<S, false>(C, cx, cy, cz, (int)(cell & 255u), cell >> 24, base + s_tri[e], verts, cap, s_val,
                                (int)threadIdx.x, edges, s_cnt);
        }
        for (unsigned int e = threadIdx.x + EB; e < ncells; e += EB) {
            const unsigned int cell = s_cell[e];
            int cx, cy, cz;
            kf_xyz(C.g, blk0 + ((cell >> 8) & 0xffffu), cx, cy, cz);
            emit_cell<S, true>(C, cx, cy, cz, (int)(cell & 255u), cell >> 24, base + s_tri[e], verts, cap, s_val,
                               (int)threadIdx.x, edges, s_cnt);
        }
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

// Per-block totals of FLAG_CHUNK consecutive flags.
__global__ __launch_bounds__(BLOCK) void k_flag_sums(const unsigned char *flags, int64_t n, unsigned int *tot)
{
    const int64_t base = (int64_t)blockIdx.x * FLAG_CHUNK;
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
    const int64_t base = (int64_t)blockIdx.x * FLAG_CHUNK;
    unsigned long long run = offs[blockIdx.x];
    for (int k = 0; k < ITEMS; k++) {
        const int64_t j = base + (int64_t)k * BLOCK + threadIdx.x;
        const unsigned int f = j < n ? flags[j] : 0u;
        unsigned int round_total;
        const unsigned int incl = block_scan_incl<BLOCK>(f, s, round_total);
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
    const int64_t i0 = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * CPT;
    int ci[CPT], x[CPT], y[CPT], z[CPT];
    C.template quad<CPT>(i0, ci, x, y, z);
    unsigned int cnt = 0;
#pragma unroll
    for (int c = 0; c < CPT; c++)
        cnt += ci[c] >= 0 ? c_tri_count[ci[c]] : 0u;
    unsigned int total;
    const unsigned int incl = block_scan_incl<BLOCK>(cnt, s, total);
    unsigned long long t0 = offs[blockIdx.x] + (incl - cnt);
    for (int c = 0; c < CPT; c++) {
        if (ci[c] < 0)
            continue;
        const unsigned int n = c_tri_count[ci[c]];
        for (unsigned int t = 0; t < n; t++)
            for (int v = 0; v < 3; v++) {
                const int e = c_tri_edges[ci[c]][t * 3 + v];
                tris[3 * (t0 + t) + v] = ids[E.id(x[c] + c_edge_node[e][0], y[c] + c_edge_node[e][1],
                                                  z[c] + c_edge_node[e][2], c_edge_node[e][3])];
            }
        t0 += n;
    }
}

// Floats in increasing order as unsigned keys (-0 just below +0; NaN outside [key(-inf), key(+inf)]).
uint32_t float_key(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : b | 0x80000000u;
}
float key_float(uint32_t k)
{
    const uint32_t b = (k & 0x80000000u) ? k & 0x7fffffffu : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
// The raw value's decode, in float like the kernels: raw * scale (BF16, F16) or raw (F32).
float decode_raw(float r, float scale, bool scaled)
{
    volatile float p = scaled ? r * scale : r;  // one rounded float product, as on the device
    return p;
}
// The least float r with decode(r) >= t, by bisection over the float keys: decode is
// nondecreasing in r (scale > 0) and decode(+inf) = +inf >= t for every non-NaN t.
float raw_least_ge(float t, float scale, bool scaled)
{
    uint32_t lo = float_key(-INFINITY), hi = float_key(INFINITY);
    if (decode_raw(-INFINITY, scale, scaled) >= t)
        return -INFINITY;
    while (hi - lo > 1) {  // decode(key lo) < t <= decode(key hi)
        const uint32_t mid = lo + (hi - lo) / 2;
        if (decode_raw(key_float(mid), scale, scaled) >= t)
            hi = mid;
        else
            lo = mid;
    }
    return key_float(hi);
}
// The greatest float r with decode(r) <= t (decode(-inf) = -inf <= t for every non-NaN t).
float raw_greatest_le(float t, float scale, bool scaled)
{
    uint32_t lo = float_key(-INFINITY), hi = float_key(INFINITY);
    if (decode_raw(INFINITY, scale, scaled) <= t)
        return INFINITY;
    while (hi - lo > 1) {  // decode(key lo) <= t < decode(key hi)
        const uint32_t mid = lo + (hi - lo) / 2;
        if (decode_raw(key_float(mid), scale, scaled) <= t)
            lo = mid;
        else
            hi = mid;
    }
    return key_float(lo);
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
    // least float >= iso and >= -clip, greatest float <= max_surface
    float f = (float)iso;
    C.iso_up = (double)f < iso ? std::nextafterf(f, INFINITY) : f;
    f = (float)(-v->clip);
    C.clip_up = (double)f < -v->clip ? std::nextafterf(f, INFINITY) : f;
    f = (float)v->max_surface;
    C.max_dn = (double)f > v->max_surface ? std::nextafterf(f, -INFINITY) : f;
    C.oob_f = v->oob() < iso ? -INFINITY : INFINITY;
    C.oob_bit = C.oob_f < C.iso_up ? 1u : 0u;
    const bool scaled = S::dtype != KF_F32;
    C.r_iso = raw_least_ge(C.iso_up, C.scale, scaled);
    C.r_clip = raw_least_ge(C.clip_up, C.scale, scaled);
    C.r_max = raw_greatest_le(C.max_dn, C.scale, scaled);
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

int64_t kf_mc_blocks(int64_t n) { return (n + CELLS - 1) / CELLS; }

namespace {

// Workgroups of the emit pass: it strides over the units of the blocks that hold triangles,
// with at most as many workgroups of EB threads as fit on the CUs at once (2048 threads each).
unsigned int emit_grid(const kf_context *ctx, int nb)
{
    const int64_t g = (2048 / EB) * (ctx->num_cus > 0 ? ctx->num_cus : 256);
    const int64_t nu = (int64_t)nb * UNITS_PER_BLOCK;
    return (unsigned int)(nu < g ? nu : g);
}

}  // namespace

kf_status kf_mc_launch_into(const kf_volume *v, double iso, unsigned int *unit_tot, unsigned int *tot,
                            unsigned long long *offs, unsigned int *blist, unsigned long long *total_out,
                            float *verts, uint64_t cap)
{
    const int nb = (int)kf_mc_blocks(v->n);
    hipStream_t st = v->ctx->stream;
    KF_DISPATCH_DTYPE(v->dtype, S, {
        const Cells<S> C = make_cells<S>(v, iso);
        if (v->dims[0] % 8 == 0)
            hipLaunchKernelGGL((k_mc_count<S, true>), dim3(nb), dim3(BLOCK), 0, st, C, unit_tot, tot);
        else
            hipLaunchKernelGGL((k_mc_count<S, false>), dim3(nb), dim3(BLOCK), 0, st, C, unit_tot, tot);
        hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, st, (const unsigned int *)tot, offs, nb, total_out,
                           blist);
        if (cap > 0)
            hipLaunchKernelGGL((k_mc_emit<S>), dim3(emit_grid(v->ctx, nb)), dim3(EB), 0, st, C,
                               (const unsigned long long *)offs, (const uint4 *)unit_tot, (const unsigned int *)blist,
                               verts, (unsigned long long)cap);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

namespace {

int mc_blocks(const kf_volume *v) { return (int)kf_mc_blocks(v->n); }

// Counts v's triangles into the context's cache (block offsets and unit totals on the device,
// the total on the host), unless the cache already holds this version of v at this iso value.
kf_status mc_count_cached(const kf_volume *v, double iso, uint64_t *total)
{
    kf_context *ctx = v->ctx;
    kf_mc_cache &c = ctx->mc;
    if (c.valid && c.vol == v && c.version == v->version && c.iso == iso) {
        *total = *c.total_host;
        return KF_OK;
    }
    c.valid = false;
    const int nb = mc_blocks(v);
    if ((size_t)nb > c.cap) {
        KF_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(c.unit_tot);
        (void)hipFree(c.tot);
        (void)hipFree(c.offs);
        (void)hipFree(c.blist);
        c.unit_tot = nullptr;
        c.tot = nullptr;
        c.offs = nullptr;
        c.blist = nullptr;
        c.cap = 0;
        KF_HIP(hipMalloc((void **)&c.unit_tot, sizeof(unsigned int) * UNITS_PER_BLOCK * (size_t)nb));
        KF_HIP(hipMalloc((void **)&c.tot, sizeof(unsigned int) * (size_t)nb));
        KF_HIP(hipMalloc((void **)&c.offs, sizeof(unsigned long long) * ((size_t)nb + 1)));
        KF_HIP(hipMalloc((void **)&c.blist, sizeof(unsigned int) * ((size_t)nb + 1)));
        c.cap = (size_t)nb;
    }
    if (!c.total_host)
        KF_HIP(hipHostMalloc((void **)&c.total_host, 2 * sizeof(unsigned long long), hipHostMallocDefault));
    kf_status s = kf_mc_launch_into(v, iso, c.unit_tot, c.tot, c.offs, c.blist, nullptr, nullptr, 0);
    if (s != KF_OK)
        return s;
    KF_HIP(hipMemcpyAsync(c.total_host, c.offs + nb, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    c.vol = v;
    c.version = v->version;
    c.iso = iso;
    c.valid = true;
    *total = *c.total_host;
    return KF_OK;
}

// Emits the cached count's triangles of v (mc_count_cached just ran for it) into dverts.
kf_status mc_emit_cached(const kf_volume *v, double iso, float *dverts, uint64_t cap)
{
    kf_context *ctx = v->ctx;
    const int nb = mc_blocks(v);
    KF_DISPATCH_DTYPE(v->dtype, S, {
        hipLaunchKernelGGL((k_mc_emit<S>), dim3(emit_grid(ctx, nb)), dim3(EB), 0, ctx->stream, make_cells<S>(v, iso),
                           (const unsigned long long *)ctx->mc.offs, (const uint4 *)ctx->mc.unit_tot,
                           (const unsigned int *)ctx->mc.blist, dverts, (unsigned long long)cap);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status grow_verts(float **buf, size_t *cap, uint64_t need, hipStream_t st)
{
    if (need <= *cap)
        return KF_OK;
    if (*buf) {
        KF_HIP(hipStreamSynchronize(st));
        KF_HIP(hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    KF_HIP(hipMalloc((void **)buf, sizeof(float) * 9 * (size_t)need));
    *cap = (size_t)need;
    return KF_OK;
}

}  // namespace

void kf_mc_release(kf_context *ctx)
{
    kf_mc_cache &c = ctx->mc;
    (void)hipFree(c.tot);
    (void)hipFree(c.offs);
    (void)hipFree(c.unit_tot);
    (void)hipFree(c.blist);
    (void)hipFree(c.verts);
    if (c.total_host)
        (void)hipHostFree(c.total_host);
    c = kf_mc_cache{};
}

kf_status kf_mc_run(const kf_volume *v, double iso, uint64_t *n_tris, float *verts_host, uint64_t cap)
{
    kf_context *ctx = v->ctx;
    uint64_t total = 0;
    kf_status s = mc_count_cached(v, iso, &total);
    if (s != KF_OK)
        return s;
    *n_tris = total;
    if (!verts_host || total == 0)
        return KF_OK;
    if (cap < total) {
        kf_set_error("marching cubes: capacity below triangle count");
        return KF_E_CAPACITY;
    }
    kf_mc_cache &c = ctx->mc;
    if ((s = grow_verts(&c.verts, &c.verts_cap, total, ctx->stream)) != KF_OK)
        return s;
    if ((s = mc_emit_cached(v, iso, c.verts, total)) != KF_OK)
        return s;
    KF_HIP(hipMemcpyAsync(verts_host, c.verts, sizeof(float) * 9 * (size_t)total, hipMemcpyDeviceToHost,
                          ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    return KF_OK;
}

kf_status kf_mc_to_device(const kf_volume *v, double iso, float **dverts, size_t *dverts_cap, uint64_t *n_tris)
{
    uint64_t total = 0;
    kf_status s = mc_count_cached(v, iso, &total);
    if (s != KF_OK)
        return s;
    *n_tris = total;
    if (total == 0)
        return KF_OK;
    if ((s = grow_verts(dverts, dverts_cap, total, v->ctx->stream)) != KF_OK)
        return s;
    return mc_emit_cached(v, iso, *dverts, total);
}

kf_status kf_mc_run_indexed(const kf_volume *v, double iso, uint64_t *n_verts, uint64_t *n_tris, float *verts_host,
                            float *normals_host, uint32_t *tris_host, uint64_t cap_verts, uint64_t cap_tris)
{
    kf_context *ctx = v->ctx;
    hipStream_t st = ctx->stream;
    uint64_t ntri = 0;
    kf_status status = mc_count_cached(v, iso, &ntri);
    if (status != KF_OK)
        return status;
    const unsigned long long *toffs = ctx->mc.offs;
    EdgeGrid E{(int64_t)v->dims[0] + 1, (int64_t)v->dims[1] + 1, (int64_t)v->dims[2] + 1};
    const int64_t ne = E.count();
    const int nbe = (int)((ne + FLAG_CHUNK - 1) / FLAG_CHUNK);
    // One scratch block: edge block totals (u32) and offsets (u64), edge flags.
    const size_t tot_bytes = ((size_t)nbe * 4 + 7) / 8 * 8;
    char *scr = nullptr;
    KF_HIP(hipMallocAsync((void **)&scr, tot_bytes + 8 * ((size_t)nbe + 1) + (size_t)ne, st));
    unsigned int *etot = (unsigned int *)scr;
    unsigned long long *eoffs = (unsigned long long *)(scr + tot_bytes);
    unsigned char *flags = (unsigned char *)(eoffs + nbe + 1);
    unsigned long long nvert = 0;
    unsigned long long *nvert_host = ctx->mc.total_host + 1;  // pinned
    hipError_t e = hipSuccess;
    unsigned int *ids = nullptr, *dtris = nullptr;
    float *dverts = nullptr, *dnorm = nullptr;
    KF_DISPATCH_DTYPE(v->dtype, S, {
        Cells<S> C = make_cells<S>(v, iso);
        e = hipMemsetAsync(flags, 0, (size_t)ne, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL((k_mc_edge_flags<S>), dim3((unsigned int)((v->n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                               st, C, E, flags);
            hipLaunchKernelGGL(k_flag_sums, dim3(nbe), dim3(BLOCK), 0, st, (const unsigned char *)flags, ne, etot);
            hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, st, (const unsigned int *)etot, eoffs, nbe,
                               (unsigned long long *)nullptr, (unsigned int *)nullptr);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(nvert_host, eoffs + nbe, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            kf_set_error(std::string("marching cubes (indexed) count: ") + hipGetErrorString(e));
            status = KF_E_HIP;
        }
        nvert = *nvert_host;
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
                    hipLaunchKernelGGL((k_mc_emit_indexed<S>), dim3(mc_blocks(v)), dim3(BLOCK), 0, st, C, E, toffs,
                                       (const unsigned int *)ids, dtris);
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
    (void)hipFreeAsync(scr, st);
    return status;
}
