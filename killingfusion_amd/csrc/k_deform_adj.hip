// k_deform_adj.hip -- fused forward-energy + adjoint-gradient + Jacobi-update kernel.
//
// Discrete energy (no reference counterpart; its parts restate the reference's terms):
//   q_j   = j + u_j (index space), psi_j = T(q_j) with T the trilinear interpolant of the
//           live SDF under the reference's read convention (SDF.cpp:503-524: truncating base,
//           out-of-grid corners = Max + eps); g_j = grad T(q_j), H_j = its mixed second
//           derivatives (a trilinear has none on the diagonal).
//   m_j   = active mask, -clip <= psi_j <= Max - eps (KillingFusion.cpp:272-274).
//   E     = sum_j m_j [ (psi_j - phi_can_j)^2 / (2 V)            data  (KillingFusion.cpp:444-455)
//                     + omega_s (s_j - 1)^2 / 2 ]                 level set (KillingFusion.cpp:464-476)
//         + omega_k sum_{i in [-1,n-1]^3} ||D u_i||_F^2 + gamma tr((D u_i)^2)
//                                                                  Killing (DisplacementField.cpp:171-193)
//   s_j = sqrt(|g_j|^2 + eps^2), (D u_i)_ab = u^a_{i+e_b} - u^a_i, u = 0 outside the grid.
// dE/du_j = m_j [ (psi_j - phi_can_j)/V g_j + omega_s (s_j - 1)/s_j H_j g_j ]
//         - 2 omega_k [ Lap u_j + gamma sum_b (D_a u^b_j - D_a u^b_{j-e_b}) ]_a
//
// One pass per iteration reads u_j, phi_can_j, the 8 live-SDF corners around q_j and the
// 1-neighbourhood of u (12 neighbours), and writes u_j - alpha dE/du_j for active voxels.
// Compulsory HBM traffic per voxel-iteration: 2 s_phi + 24 bytes.
//
// Two launch shapes share the per-voxel arithmetic (adj_voxel):
//  * k_adj_rows (rows whose length is a multiple of 4): a thread owns 4 consecutive x voxels
//    of two rows, a wavefront 256 voxels of a row, a block 8 rows of one slice.  A Jacobi step
//    moves only active voxels: a voxel whose own u is zero samples phi_live[j] exactly, so the
//    rows of u, phi_live and u_out stream as float4 / 8-byte accesses and only active voxels
//    (and moved ones, to find out) run the stencil.  Energy and gradient passes, which need
//    every voxel whose stencil sees a nonzero u, take that set from a bit mask of the 4-voxel
//    groups of u that may be nonzero.
//  * k_adj_any (any shape): one voxel per thread, neighbours read individually.
// Blocks are numbered so each XCD sweeps a contiguous z range and the rows and slices shared
// by neighbouring blocks stay in that XCD's L2.
#include "kf_reduce.hpp"

namespace {

constexpr int BLOCK = 256;
constexpr int VEC = 4;                  // voxels per thread along x in k_adj_rows
constexpr int ROW = 64 * VEC;           // voxels of a row segment per wavefront
constexpr int RPT = 2;                  // rows per thread in k_adj_rows (y and y + 4)
constexpr int ROWS = BLOCK / 64 * RPT;  // rows per block (a row tile)

struct AdjConsts {
    float inv_2v, inv_v, hi, lo, eps2, alpha, omega_k, gamma, omega_s;
    int use_data, use_ls, use_k;
};

template <class S>
struct PhiRead {
    const typename S::T *p;
    GridDims g;
    float scale;
    float oob;
    __device__ inline float at(int x, int y, int z) const
    {
        return g.in(x, y, z) ? S::load(p, g.idx(x, y, z), scale) : oob;
    }
    // The 8 corners of cell (bx, by, bz) in utils.h order (x fastest); out-of-grid corners
    // read Max + eps (SDF.cpp:482-487).
    __device__ inline void cell(int bx, int by, int bz, float v[8]) const
    {
        if (bx >= 0 && by >= 0 && bz >= 0 && bx + 1 < g.nx && by + 1 < g.ny && bz + 1 < g.nz) {
            const int64_t i0 = g.idx(bx, by, bz);
            const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
            v[0] = S::load(p, i0, scale);
            v[1] = S::load(p, i0 + 1, scale);
            v[2] = S::load(p, i0 + sy, scale);
            v[3] = S::load(p, i0 + sy + 1, scale);
            v[4] = S::load(p, i0 + sz, scale);
            v[5] = S::load(p, i0 + sz + 1, scale);
            v[6] = S::load(p, i0 + sz + sy, scale);
            v[7] = S::load(p, i0 + sz + sy + 1, scale);
        } else {
            v[0] = at(bx, by, bz);
            v[1] = at(bx + 1, by, bz);
            v[2] = at(bx, by + 1, bz);
            v[3] = at(bx + 1, by + 1, bz);
            v[4] = at(bx, by, bz + 1);
            v[5] = at(bx + 1, by, bz + 1);
            v[6] = at(bx, by + 1, bz + 1);
            v[7] = at(bx + 1, by + 1, bz + 1);
        }
    }
};

// u neighbourhood of voxel j: centre, face neighbours up[b][a] = u^a_{j+e_b},
// um[b][a] = u^a_{j-e_b}, and the six edge neighbours dgl[a][b] = u^b_{j+e_a-e_b} (a != b)
// the gamma term needs.  Zero outside the grid.
struct Nbhd {
    float uc[3];
    float up[3][3], um[3][3];
    float dgl[3][3];
};

struct VoxelOut {
    float g[3];
    float e;
    bool active;
};

// dE/du_j and the energy owned by voxel j (see the file header for E).  `v` are the 8
// live-SDF corners of the cell holding j + u_j, (fx, fy, fz) its weights.
template <bool NEED_E>
__device__ inline void adj_voxel(const AdjConsts &c, const float v[8], float fx, float fy, float fz, float can,
                                 const Nbhd &nb, int x, int y, int z, VoxelOut &o)
{
    const float psi = kf_interp3(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], fx, fy, fz);
    o.active = !(psi > c.hi || psi < c.lo);
    float gr0 = 0.f, gr1 = 0.f, gr2 = 0.f, e = 0.f;
    if (o.active) {
        const float dx00 = v[1] - v[0], dx10 = v[3] - v[2], dx01 = v[5] - v[4], dx11 = v[7] - v[6];
        const float dy00 = v[2] - v[0], dy10 = v[3] - v[1], dy01 = v[6] - v[4], dy11 = v[7] - v[5];
        const float dz00 = v[4] - v[0], dz10 = v[5] - v[1], dz01 = v[6] - v[2], dz11 = v[7] - v[3];
        const float g0 = kf_lerp(kf_lerp(dx00, dx10, fy), kf_lerp(dx01, dx11, fy), fz);
        const float g1 = kf_lerp(kf_lerp(dy00, dy10, fx), kf_lerp(dy01, dy11, fx), fz);
        const float g2 = kf_lerp(kf_lerp(dz00, dz10, fx), kf_lerp(dz01, dz11, fx), fy);
        if (c.use_data) {
            const float res = psi - can;
            if (NEED_E)
                e += res * res * c.inv_2v;
            const float s = res * c.inv_v;
            gr0 += s * g0;
            gr1 += s * g1;
            gr2 += s * g2;
        }
        if (c.use_ls) {
            const float hxy = kf_lerp(dx10 - dx00, dx11 - dx01, fz);
            const float hxz = kf_lerp(dx01 - dx00, dx11 - dx10, fy);
            const float hyz = kf_lerp(dy01 - dy00, dy11 - dy10, fx);
            const float sl = sqrtf(g0 * g0 + g1 * g1 + g2 * g2 + c.eps2);
            if (NEED_E)
                e += c.omega_s * 0.5f * (sl - 1.0f) * (sl - 1.0f);
            const float k = c.omega_s * (sl - 1.0f) / sl;
            gr0 += k * (hxy * g1 + hxz * g2);
            gr1 += k * (hxy * g0 + hyz * g2);
            gr2 += k * (hxz * g0 + hyz * g1);
        }
    }
    if (c.use_k) {
        float kg[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float lap = 0.f;
#pragma unroll
            for (int b = 0; b < 3; b++)
                lap += nb.up[b][a] - 2.0f * nb.uc[a] + nb.um[b][a];
            float dv = 0.f;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                // D_a u^b at j and at j - e_b
                const float d_j = nb.up[a][b] - nb.uc[b];
                const float d_m = (a == b) ? (nb.uc[b] - nb.um[b][b]) : (nb.dgl[a][b] - nb.um[b][b]);
                dv += d_j - d_m;
            }
            kg[a] = -2.0f * (lap + c.gamma * dv);
        }
        gr0 += c.omega_k * kg[0];
        gr1 += c.omega_k * kg[1];
        gr2 += c.omega_k * kg[2];
        if (NEED_E) {
            // ||D u_j||^2 + gamma tr((D u_j)^2) + the ghost terms of the i = j - e_c voxels
            // on low faces.
            float ek = 0.f, tr = 0.f;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    const float dab = nb.up[b][a] - nb.uc[a];
                    const float dba = nb.up[a][b] - nb.uc[b];
                    ek += dab * dab;
                    tr += dab * dba;
                }
            ek += c.gamma * tr;
            const int jj[3] = {x, y, z};
            const float s2 = nb.uc[0] * nb.uc[0] + nb.uc[1] * nb.uc[1] + nb.uc[2] * nb.uc[2];
#pragma unroll
            for (int cc = 0; cc < 3; cc++)
                if (jj[cc] == 0)
                    ek += s2 + c.gamma * nb.uc[cc] * nb.uc[cc];
            e += c.omega_k * ek;
        }
    }
    o.g[0] = gr0;
    o.g[1] = gr1;
    o.g[2] = gr2;
    o.e = e;
}

// Jacobi update of one voxel: active voxels move by -alpha dE/du; non-finite results are
// reported (KillingFusion.cpp:327-332).  Returns the new value in n[3].
__device__ inline void adj_update(const AdjConsts &c, const VoxelOut &o, const float uc[3], int64_t i,
                                  long long *div, double &mx, float n[3])
{
    n[0] = uc[0];
    n[1] = uc[1];
    n[2] = uc[2];
    if (o.active) {
        const float d0 = -c.alpha * o.g[0], d1 = -c.alpha * o.g[1], d2 = -c.alpha * o.g[2];
        n[0] = uc[0] + d0;
        n[1] = uc[1] + d1;
        n[2] = uc[2] + d2;
        if (!(isfinite(n[0]) && isfinite(n[1]) && isfinite(n[2])))
            atomicMin(div, (long long)i);
        else
            mx = fmax(mx, sqrt((double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2));
    }
}

__device__ inline float uval(const float *__restrict__ u, const GridDims &g, int c, int x, int y, int z)
{
    return g.in(x, y, z) ? u[(int64_t)c * g.n + g.idx(x, y, z)] : 0.0f;
}

// Gate of step launch k (parity k & 1): the launch runs iff st->run[parity].  Its first block
// folds launch k - 1 (prev_nb partials) and then arms run[parity ^ 1] for launch k + 1, so
// a launch never reads a gate field it writes: every wave of every block sees the same
// answer.  Launch k itself may run after launch k - 1 met the stop rule; it then only writes
// the buffer that is not the result (see kf_deform), and launch k + 1 does nothing.
template <int MODE>
__device__ inline bool step_gate(DeformState *st, const BlockPartial *prev, unsigned int prev_nb, int parity)
{
    if (MODE != KF_MODE_STEP)
        return true;
    const bool go = st->run[parity] != 0u;
    if (blockIdx.x == 0) {
        if (go && prev_nb > 0) {
            kf_fold_launch<BLOCK>(st, prev, prev_nb, parity ^ 1);
            __syncthreads();  // the block's own publish reuses the reduction LDS
        }
        if (threadIdx.x == 0)
            st->run[parity ^ 1] = (go && !st->done) ? 1u : 0u;
    }
    return go;
}

// ---------------------------------------------------------------------------
// Any shape: one voxel per thread.
// ---------------------------------------------------------------------------
template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void k_adj_any(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                   float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                   BlockPartial *partials, float *__restrict__ e_voxel, int parity,
                                                   const BlockPartial *prev, unsigned int prev_nb)
{
    if (!step_gate<MODE>(st, prev, prev_nb, parity))
        return;
    const GridDims g = L.g;
    const int64_t i = (int64_t)kf_xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    if (i < g.n) {
        const int x = (int)(i % g.nx);
        const int y = (int)((i / g.nx) % g.ny);
        const int z = (int)(i / ((int64_t)g.nx * g.ny));
        Nbhd nb;
        nb.uc[0] = u_in[i];
        nb.uc[1] = u_in[g.n + i];
        nb.uc[2] = u_in[2 * g.n + i];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            nb.up[0][a] = uval(u_in, g, a, x + 1, y, z);
            nb.um[0][a] = uval(u_in, g, a, x - 1, y, z);
            nb.up[1][a] = uval(u_in, g, a, x, y + 1, z);
            nb.um[1][a] = uval(u_in, g, a, x, y - 1, z);
            nb.up[2][a] = uval(u_in, g, a, x, y, z + 1);
            nb.um[2][a] = uval(u_in, g, a, x, y, z - 1);
        }
        nb.dgl[0][1] = uval(u_in, g, 1, x + 1, y - 1, z);
        nb.dgl[0][2] = uval(u_in, g, 2, x + 1, y, z - 1);
        nb.dgl[1][0] = uval(u_in, g, 0, x - 1, y + 1, z);
        nb.dgl[1][2] = uval(u_in, g, 2, x, y + 1, z - 1);
        nb.dgl[2][0] = uval(u_in, g, 0, x - 1, y, z + 1);
        nb.dgl[2][1] = uval(u_in, g, 1, x, y - 1, z + 1);
        int bx, by, bz;
        float fx, fy, fz, v[8];
        kf_split(x, nb.uc[0], bx, fx);
        kf_split(y, nb.uc[1], by, fy);
        kf_split(z, nb.uc[2], bz, fz);
        L.cell(bx, by, bz, v);
        VoxelOut o;
        adj_voxel<MODE == KF_MODE_ENERGY>(c, v, fx, fy, fz, c.use_data ? Cn.at(x, y, z) : 0.f, nb, x, y, z, o);
        if (o.active)
            act = 1;
        if (MODE == KF_MODE_GRAD) {
            u_out[i] = o.g[0];
            u_out[g.n + i] = o.g[1];
            u_out[2 * g.n + i] = o.g[2];
        } else if (MODE == KF_MODE_ENERGY) {
            if (e_voxel)
                e_voxel[i] = o.e;
        } else {
            float n[3];
            adj_update(c, o, nb.uc, i, &st->div_slot[parity], mx, n);
            u_out[i] = n[0];
            u_out[g.n + i] = n[1];
            u_out[2 * g.n + i] = n[2];
        }
        esum = (double)o.e;
    }
    if (MODE != KF_MODE_GRAD)
        kf_block_publish<BLOCK>(partials, blockIdx.x, esum, mx, act);
}

// ---------------------------------------------------------------------------
// Rows of a multiple of 4 voxels: 4 voxels per thread, float4 row loads.
// ---------------------------------------------------------------------------

// The row kernel indexes with 32-bit element offsets from the kernel-argument base
// pointers (3 N < 2^30 is checked at launch), which keeps one VGPR per address.
__device__ inline unsigned int nvox32(const GridDims &g) { return (unsigned int)g.n; }

// Row segment of component c at (x0 .. x0+3, y, z); zero when the row is out of the grid.
__device__ inline float4 row4(const float *__restrict__ u, const GridDims &g, int c, int x0, int y, int z)
{
    if (x0 < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz)
        return *reinterpret_cast<const float4 *>(u + ((unsigned int)c * nvox32(g) + g.idx32(x0, y, z)));
    return make_float4(0.f, 0.f, 0.f, 0.f);
}

__device__ inline float uval32(const float *__restrict__ u, const GridDims &g, int c, int x, int y, int z)
{
    return g.in(x, y, z) ? u[(unsigned int)c * nvox32(g) + g.idx32(x, y, z)] : 0.0f;
}

template <class S>
__device__ inline float at32(const PhiRead<S> &P, int x, int y, int z)
{
    return P.g.in(x, y, z) ? S::load(P.p, (int64_t)P.g.idx32(x, y, z), P.scale) : P.oob;
}

// PhiRead::cell with 32-bit offsets.
template <class S>
__device__ inline void cell32(const PhiRead<S> &P, int bx, int by, int bz, float v[8])
{
    const GridDims &g = P.g;
    if (bx >= 0 && by >= 0 && bz >= 0 && bx + 1 < g.nx && by + 1 < g.ny && bz + 1 < g.nz) {
        const typename S::T *p0 = P.p + g.idx32(bx, by, bz);
        const unsigned int sy = (unsigned int)g.nx, sz = (unsigned int)g.nx * (unsigned int)g.ny;
        v[0] = S::load(p0, 0, P.scale);
        v[1] = S::load(p0, 1, P.scale);
        v[2] = S::load(p0, sy, P.scale);
        v[3] = S::load(p0, sy + 1, P.scale);
        v[4] = S::load(p0, sz, P.scale);
        v[5] = S::load(p0, sz + 1, P.scale);
        v[6] = S::load(p0, sz + sy, P.scale);
        v[7] = S::load(p0, sz + sy + 1, P.scale);
    } else {
        v[0] = at32(P, bx, by, bz);
        v[1] = at32(P, bx + 1, by, bz);
        v[2] = at32(P, bx, by + 1, bz);
        v[3] = at32(P, bx + 1, by + 1, bz);
        v[4] = at32(P, bx, by, bz + 1);
        v[5] = at32(P, bx + 1, by, bz + 1);
        v[6] = at32(P, bx, by + 1, bz + 1);
        v[7] = at32(P, bx + 1, by + 1, bz + 1);
    }
}

__device__ inline unsigned int bits4(const float4 &v)
{
    return __float_as_uint(v.x) | __float_as_uint(v.y) | __float_as_uint(v.z) | __float_as_uint(v.w);
}

__device__ inline float f4(const float4 &v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// u neighbourhood of voxel (x, y, z) read value by value (cache hits: the rows were just
// streamed in by the quiet test).
__device__ inline void gather_nbhd(const float *__restrict__ u, const GridDims &g, int x, int y, int z, Nbhd &nb)
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        nb.uc[a] = uval32(u, g, a, x, y, z);
        nb.up[0][a] = uval32(u, g, a, x + 1, y, z);
        nb.um[0][a] = uval32(u, g, a, x - 1, y, z);
        nb.up[1][a] = uval32(u, g, a, x, y + 1, z);
        nb.um[1][a] = uval32(u, g, a, x, y - 1, z);
        nb.up[2][a] = uval32(u, g, a, x, y, z + 1);
        nb.um[2][a] = uval32(u, g, a, x, y, z - 1);
    }
    nb.dgl[0][1] = uval32(u, g, 1, x + 1, y - 1, z);
    nb.dgl[0][2] = uval32(u, g, 2, x + 1, y, z - 1);
    nb.dgl[1][0] = uval32(u, g, 0, x - 1, y + 1, z);
    nb.dgl[1][2] = uval32(u, g, 2, x, y + 1, z - 1);
    nb.dgl[2][0] = uval32(u, g, 0, x - 1, y, z + 1);
    nb.dgl[2][1] = uval32(u, g, 1, x, y - 1, z + 1);
}

// Nonzero masks.  mask bit (row, group) = 1 iff some u component of the 4-voxel group
// (x = 4 group .. 4 group + 3) of that row may be nonzero.  A row (y, z) holds MW words per
// 256-voxel segment (one word per half wavefront), so each wavefront owns its two words and
// writes them with a ballot -- no atomics.
constexpr int MW = ROW / VEC / 32;

// Bits of groups gx - 1 .. gx + 1 of row (y, z) (gx = 64 seg + lane); 0 outside the grid.
__device__ inline unsigned int nbr_bits(const unsigned int *__restrict__ m, int wpr, const GridDims &g, int y, int z,
                                        int seg, int lane)
{
    if (y < 0 || y >= g.ny || z < 0 || z >= g.nz)
        return 0u;
    const unsigned int *row = m + ((unsigned int)z * (unsigned int)g.ny + (unsigned int)y) * (unsigned int)wpr;
    const int gx = seg * (ROW / VEC) + lane;
    const int wi = gx >> 5, b = gx & 31;
    unsigned int r = row[wi] & (b == 0 ? 0x3u : (0x7u << (b - 1)));
    if (b == 0 && wi > 0)
        r |= row[wi - 1] >> 31;
    if (b == 31 && wi + 1 < wpr)
        r |= row[wi + 1] & 1u;
    return r;
}

// Writes the wavefront's mask words of row (y, z): bit lane = nz of that lane's group.
__device__ inline void put_bits(unsigned int *__restrict__ m, int wpr, const GridDims &g, int y, int z, int seg,
                                int lane, bool nz)
{
    const unsigned long long bal = __ballot(nz);
    if (lane == 0 && y < g.ny) {
        unsigned int *row = m + ((unsigned int)z * (unsigned int)g.ny + (unsigned int)y) * (unsigned int)wpr;
        row[seg * MW] = (unsigned int)bal;
        row[seg * MW + 1] = (unsigned int)(bal >> 32);
    }
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void k_adj_rows(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                    BlockPartial *partials, float *__restrict__ e_voxel, int nseg,
                                                    int nyt, const unsigned int *__restrict__ mask_in, int parity,
                                                    const BlockPartial *prev, unsigned int prev_nb)
{
    if (!step_gate<MODE>(st, prev, prev_nb, parity))
        return;
    const GridDims g = L.g;
    const unsigned int tile = kf_xcd_tile(blockIdx.x, gridDim.x);
    const int seg = (int)(tile % (unsigned)nseg);
    const int yt = (int)((tile / (unsigned)nseg) % (unsigned)nyt);
    const int z = (int)(tile / ((unsigned)nseg * (unsigned)nyt));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x0 = seg * ROW + lane * VEC;
    const int wpr = nseg * MW;
    const unsigned int n32 = nvox32(g);
    int yr[RPT];
    bool valid[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        yr[r] = yt * ROWS + w + r * (BLOCK / 64);
        valid[r] = x0 < g.nx && yr[r] < g.ny;
    }

    // Phase 1 -- quiet test.  A voxel is quiet when every u value its stencil touches is
    // zero (everywhere away from the surface band); then psi_j = phi_live[j] exactly (zero
    // interpolation weights) and an inactive quiet voxel has zero gradient and energy.  The
    // stencil of a thread's 4 voxels reaches groups gx - 1 .. gx + 1 of seven rows; their mask
    // bits decide quietness without touching u.  Both rows' loads are issued up front.
    float4 C[RPT][3];
    float live4[RPT][VEC];
    unsigned int nb_bits[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int y = yr[r];
#pragma unroll
        for (int a = 0; a < 3; a++)
            C[r][a] = row4(u_in, g, a, x0, y, z);
#pragma unroll
        for (int k = 0; k < VEC; k++)
            live4[r][k] = 0.f;
        if (valid[r])
            S::load4(L.p, (int64_t)g.idx32(x0, y, z), L.scale, live4[r]);
        nb_bits[r] = 0u;
        if (MODE != KF_MODE_STEP)
            nb_bits[r] = nbr_bits(mask_in, wpr, g, y, z, seg, lane) | nbr_bits(mask_in, wpr, g, y + 1, z, seg, lane) |
                         nbr_bits(mask_in, wpr, g, y - 1, z, seg, lane) |
                         nbr_bits(mask_in, wpr, g, y, z + 1, seg, lane) |
                         nbr_bits(mask_in, wpr, g, y, z - 1, seg, lane) |
                         nbr_bits(mask_in, wpr, g, y + 1, z - 1, seg, lane) |
                         nbr_bits(mask_in, wpr, g, y - 1, z + 1, seg, lane);
    }
    // need bit r * VEC + k: voxel (x0 + k, yr[r]) goes to phase 2.  A Jacobi step moves only
    // active voxels, so it needs the voxels that are active or whose sample moved (u_j != 0,
    // activity unknown until the gather); energy and gradient passes also need every voxel
    // whose stencil sees a nonzero u.
    unsigned int need = 0u;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const bool quiet = nb_bits[r] == 0u && ((bits4(C[r][0]) | bits4(C[r][1]) | bits4(C[r][2])) & 0x7fffffffu) == 0u;
#pragma unroll
        for (int k = 0; k < VEC; k++) {
            const bool moved = ((__float_as_uint(f4(C[r][0], k)) | __float_as_uint(f4(C[r][1], k)) |
                                 __float_as_uint(f4(C[r][2], k))) & 0x7fffffffu) != 0u;
            const bool in_band = !(live4[r][k] > c.hi || live4[r][k] < c.lo);
            if (valid[r] && (MODE == KF_MODE_STEP ? (moved || in_band) : (!quiet || in_band)))
                need |= 1u << (r * VEC + k);
        }
    }

    // Defaults of the voxels that cannot contribute, four at a time; phase-2 voxels overwrite
    // their own elements below (same thread, so program order holds).
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        if (!valid[r])
            continue;
        const unsigned int i0 = g.idx32(x0, yr[r], z);
        if (MODE == KF_MODE_ENERGY) {
            if (e_voxel)
                *reinterpret_cast<float4 *>(e_voxel + i0) = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float4 d = MODE == KF_MODE_GRAD ? make_float4(0.f, 0.f, 0.f, 0.f) : C[r][a];
                *reinterpret_cast<float4 *>(u_out + ((unsigned int)a * n32 + i0)) = d;
            }
        }
    }

    // Phase 2 -- the voxels that can contribute: full energy and gradient.  The wavefront's
    // need bits become one job list (ordered by bit, then lane), so a band segment held by one
    // or two lanes is spread over the wavefront: a lane runs one gather chain instead of up to
    // eight in a row.  A job's stores may hit elements another lane of this wavefront stored
    // as defaults above; a wavefront's vector memory operations are performed in order.
    unsigned long long bal[RPT * VEC];
    int total = 0;
#pragma unroll
    for (int b = 0; b < RPT * VEC; b++) {
        bal[b] = __ballot((need >> b) & 1u);
        total += __popcll(bal[b]);
    }
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    for (int j = lane; j < total; j += 64) {
        // Job j -> need bit `bit` of lane `src`: skip whole bits, then find the rem-th set
        // lane of that bit's ballot by halving.
        int rem = j, bit = 0;
        unsigned long long m = 0ull;
        bool found = false;
#pragma unroll
        for (int b = 0; b < RPT * VEC; b++) {
            const int cnt = __popcll(bal[b]);
            if (!found) {
                if (rem < cnt) {
                    bit = b;
                    m = bal[b];
                    found = true;
                } else {
                    rem -= cnt;
                }
            }
        }
        int src = 0;
#pragma unroll
        for (int h = 32; h >= 1; h >>= 1) {
            const int lo = __popcll(m & ((1ull << h) - 1ull));
            if (rem >= lo) {
                rem -= lo;
                m >>= h;
                src += h;
            }
        }
        const int r = bit / VEC;
        const int x = seg * ROW + src * VEC + bit % VEC;
        const int y = yt * ROWS + w + r * (BLOCK / 64);
        if (!(x < g.nx && y < g.ny))
            continue;
        const unsigned int i = g.idx32(x, y, z);
        int bx, by, bz;
        float fx, fy, fz, v[8];
        {
            const float u0 = u_in[i], u1 = u_in[n32 + i], u2 = u_in[2u * n32 + i];
            kf_split(x, u0, bx, fx);
            kf_split(y, u1, by, fy);
            kf_split(z, u2, bz, fz);
        }
        // The cell corners and the u neighbourhood load together (one round trip); an
        // inactive voxel of a step keeps the u_j already stored.
        cell32(L, bx, by, bz, v);
        Nbhd nb;
        gather_nbhd(u_in, g, x, y, z, nb);
        VoxelOut o;
        adj_voxel<MODE == KF_MODE_ENERGY>(c, v, fx, fy, fz, c.use_data ? S::load(Cn.p, (int64_t)i, Cn.scale) : 0.f,
                                          nb, x, y, z, o);
        if (o.active)
            act += 1;
        esum += (double)o.e;
        if (MODE == KF_MODE_GRAD) {
            u_out[i] = o.g[0];
            u_out[n32 + i] = o.g[1];
            u_out[2u * n32 + i] = o.g[2];
        } else if (MODE == KF_MODE_ENERGY) {
            if (e_voxel)
                e_voxel[i] = o.e;
        } else if (o.active) {
            float n[3];
            adj_update(c, o, nb.uc, (int64_t)i, &st->div_slot[parity], mx, n);
            u_out[i] = n[0];
            u_out[n32 + i] = n[1];
            u_out[2u * n32 + i] = n[2];
        }
    }
    if (MODE != KF_MODE_GRAD)
        kf_block_publish<BLOCK>(partials, blockIdx.x, esum, mx, act);
}

// Nonzero mask of field values u (see nbr_bits).
__global__ __launch_bounds__(BLOCK) void k_adj_mask(const float *__restrict__ u, GridDims g, int nseg, int nyt,
                                                    unsigned int *__restrict__ mask)
{
    const unsigned int tile = blockIdx.x;
    const int seg = (int)(tile % (unsigned)nseg);
    const int yt = (int)((tile / (unsigned)nseg) % (unsigned)nyt);
    const int z = (int)(tile / ((unsigned)nseg * (unsigned)nyt));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x0 = seg * ROW + lane * VEC;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int y = yt * ROWS + w + r * (BLOCK / 64);
        unsigned int b = 0u;
#pragma unroll
        for (int a = 0; a < 3; a++)
            b |= bits4(row4(u, g, a, x0, y, z));
        put_bits(mask, nseg * MW, g, y, z, seg, lane, (b & 0x7fffffffu) != 0u);
    }
}

template <class S>
PhiRead<S> make_read(const kf_volume *v)
{
    PhiRead<S> P;
    P.p = (const typename S::T *)v->phi;
    P.g = grid_of(v);
    P.scale = v->scale();
    P.oob = (float)v->oob();
    return P;
}

template <class S, int MODE>
void launch(const PhiRead<S> &L, const PhiRead<S> &Cn, const float *u_in, float *u_out, const AdjConsts &c,
            kf_context *ctx, float *e_voxel, bool rows, unsigned int nb, int nseg, int nyt,
            const unsigned int *mask_in, int parity, const BlockPartial *prev, unsigned int prev_nb)
{
    BlockPartial *parts = kf_partials(ctx, parity);
    if (rows)
        hipLaunchKernelGGL((k_adj_rows<S, MODE>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in, u_out, c,
                           ctx->state, parts, e_voxel, nseg, nyt, mask_in, parity, prev, prev_nb);
    else
        hipLaunchKernelGGL((k_adj_any<S, MODE>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in, u_out, c,
                           ctx->state, parts, e_voxel, parity, prev, prev_nb);
}

}  // namespace

size_t kf_adj_mask_words(const int32_t dims[3])
{
    if (dims[0] % VEC != 0 || 3 * (int64_t)dims[0] * dims[1] * dims[2] >= ((int64_t)1 << 30))
        return 0;
    return (size_t)((dims[0] + ROW - 1) / ROW) * MW * (size_t)dims[1] * (size_t)dims[2];
}

kf_status kf_launch_adj_mask(const kf_field *f, const float *u, unsigned int *mask)
{
    const int nseg = (f->dims[0] + ROW - 1) / ROW, nyt = (f->dims[1] + ROWS - 1) / ROWS;
    KF_CHECK_ARG(kf_adj_mask_words(f->dims) > 0, "nonzero masks need rows of a multiple of 4 voxels");
    const unsigned int nb = (unsigned int)(nseg * nyt * f->dims[2]);
    hipLaunchKernelGGL(k_adj_mask, dim3(nb), dim3(BLOCK), 0, f->ctx->stream, u, grid_of(f), nseg, nyt, mask);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_adj_step(const kf_volume *live, const kf_volume *can, const float *u_in, float *u_out,
                             const kf_params *p, int mode, float *e_voxel, const unsigned int *mask_in, int iter,
                             unsigned int *nb_out)
{
    kf_context *ctx = live->ctx;
    AdjConsts c{};
    c.inv_2v = (float)(1.0 / (2.0 * p->voxel_size));
    c.inv_v = (float)(1.0 / p->voxel_size);
    c.hi = (float)(p->max_surface - p->epsilon);
    c.lo = (float)(-p->clip_distance);
    c.eps2 = (float)(p->epsilon * p->epsilon);
    c.alpha = (float)p->alpha;
    c.omega_k = (float)p->omega_killing;
    c.gamma = (float)p->gamma_killing;
    c.omega_s = (float)p->omega_levelset;
    c.use_data = p->use_data;
    c.use_ls = p->use_levelset;
    c.use_k = p->use_killing;
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    // The row kernel needs rows of a multiple of 4 voxels (float4 loads, 16-byte aligned plane
    // starts), 3 N < 2^30 (32-bit offsets) and, for energy and gradient passes, the nonzero mask
    // of u_in.
    const bool rows = kf_adj_mask_words(live->dims) > 0 && (mode == KF_MODE_STEP || mask_in != nullptr);
    const int nseg = (live->dims[0] + ROW - 1) / ROW;
    const int nyt = (live->dims[1] + ROWS - 1) / ROWS;
    const unsigned int nb = rows ? (unsigned int)(nseg * nyt * live->dims[2])
                                 : (unsigned int)((live->n + BLOCK - 1) / BLOCK);
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    if (nb_out)
        *nb_out = nb;
    // Energy and gradient passes use partial set 0; step launch k uses set k & 1 and folds
    // launch k - 1 (same shape: the dims do not change within a deform call).
    const int parity = mode == KF_MODE_STEP ? (iter & 1) : 0;
    const BlockPartial *prev = kf_partials(ctx, parity ^ 1);
    const unsigned int prev_nb = mode == KF_MODE_STEP && iter > 0 ? nb : 0u;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        PhiRead<S> L = make_read<S>(live), Cn = make_read<S>(can);
        if (mode == KF_MODE_GRAD)
            launch<S, KF_MODE_GRAD>(L, Cn, u_in, u_out, c, ctx, e_voxel, rows, nb, nseg, nyt, mask_in, 0, prev, 0u);
        else if (mode == KF_MODE_ENERGY)
            launch<S, KF_MODE_ENERGY>(L, Cn, u_in, u_out, c, ctx, e_voxel, rows, nb, nseg, nyt, mask_in, 0, prev,
                                      0u);
        else
            launch<S, KF_MODE_STEP>(L, Cn, u_in, u_out, c, ctx, e_voxel, rows, nb, nseg, nyt, nullptr, parity, prev,
                                    prev_nb);
    });
    KF_LAUNCH_CHECK();
    return mode == KF_MODE_ENERGY ? kf_launch_finalize(ctx, nb, 0) : KF_OK;
}
