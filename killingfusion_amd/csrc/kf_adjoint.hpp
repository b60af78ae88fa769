// kf_adjoint.hpp -- per-voxel arithmetic of the adjoint formulation, shared by the single-
// volume kernels (k_deform_adj.hip) and the batched persistent kernel (k_deform_batch.hip).
// The discrete energy and its gradient are stated at the top of k_deform_adj.hip.
#pragma once

#include "kf_reduce.hpp"

namespace {

constexpr int BLOCK = 256;

// Loads of u that stream through once per iteration carry the non-temporal hint: the lines
// are evicted first, which keeps the caches for the buffer this iteration writes (the next
// iteration reads it back) and for the SDFs.  Measured at 256^3: a step with these hints
// runs in ~75% of the time of the same step with plain loads; the hint on the stores or on
// the SDF loads made it slower.
typedef float kf_f4v __attribute__((ext_vector_type(4)));
__device__ inline float4 kf_load_stream4(const float *p)
{
    const kf_f4v v = __builtin_nontemporal_load(reinterpret_cast<const kf_f4v *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline float kf_load_stream(const float *p) { return __builtin_nontemporal_load(p); }
// The same for a group that starts on a 4-byte boundary only (one 16-byte load on gfx950).
typedef float kf_f4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ inline float4 kf_load_stream4u(const float *p)
{
    const kf_f4u v = __builtin_nontemporal_load(reinterpret_cast<const kf_f4u *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

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
template <int MODE, int NT = BLOCK>
__device__ inline bool step_gate(DeformState *st, const BlockPartial *prev, unsigned int prev_nb, int parity)
{
    if (MODE != KF_MODE_STEP)
        return true;
    const bool go = st->run[parity] != 0u;
    if (blockIdx.x == 0) {
        if (go && prev_nb > 0) {
            kf_fold_launch<NT>(st, prev, prev_nb, parity ^ 1);
            __syncthreads();  // the block's own publish reuses the reduction LDS
        }
        if (threadIdx.x == 0)
            st->run[parity ^ 1] = (go && !st->done) ? 1u : 0u;
    }
    return go;
}

// The row kernel indexes with 32-bit element offsets from the kernel-argument base
// pointers (3 N < 2^30 is checked at launch), which keeps one VGPR per address.
__device__ inline unsigned int nvox32(const GridDims &g) { return (unsigned int)g.n; }

// Row segment of component c at (x0 .. x0+3, y, z); zero when the row is out of the grid.
__device__ inline float4 row4(const float *__restrict__ u, const GridDims &g, int c, int x0, int y, int z)
{
    if (x0 < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz)
        return kf_load_stream4(u + ((unsigned int)c * nvox32(g) + g.idx32(x0, y, z)));
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


// The full stencil of voxel (x, y, z) (in the grid): u_j, the 8 live corners around
// j + u_j, the u neighbourhood and phi_can[j].  A step writes the moved u_j of an active
// voxel (the caller stored u_j before) or, with WRITE_ALL, the new u_j of every voxel; a
// gradient pass writes dE/du_j; an energy pass writes the voxel's energy when e_voxel is
// given.  Loads of the cell and the neighbourhood are issued together (one round trip after
// u_j).
template <class S, int MODE, bool WRITE_ALL = false>
__device__ inline void adj_job(const PhiRead<S> &L, const PhiRead<S> &Cn, const float *__restrict__ u_in,
                               float *__restrict__ u_out, const AdjConsts &c, long long *div,
                               float *__restrict__ e_voxel, int x, int y, int z, double &mx, double &esum,
                               unsigned long long &act)
{
    const GridDims &g = L.g;
    const unsigned int n32 = nvox32(g);
    const unsigned int i = g.idx32(x, y, z);
    int bx, by, bz;
    float fx, fy, fz, v[8];
    {
        const float u0 = u_in[i], u1 = u_in[n32 + i], u2 = u_in[2u * n32 + i];
        kf_split(x, u0, bx, fx);
        kf_split(y, u1, by, fy);
        kf_split(z, u2, bz, fz);
    }
    cell32(L, bx, by, bz, v);
    Nbhd nb;
    gather_nbhd(u_in, g, x, y, z, nb);
    VoxelOut o;
    adj_voxel<MODE == KF_MODE_ENERGY>(c, v, fx, fy, fz, c.use_data ? S::load(Cn.p, (int64_t)i, Cn.scale) : 0.f, nb,
                                      x, y, z, o);
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
    } else if (o.active || WRITE_ALL) {
        float n[3];
        adj_update(c, o, nb.uc, (int64_t)i, div, mx, n);
        u_out[i] = n[0];
        u_out[n32 + i] = n[1];
        u_out[2u * n32 + i] = n[2];
    }
}

// A wavefront's job list: lane l holds NB need bits; bal[b] is the ballot of bit b.  Jobs are
// ordered by bit, then lane.  Job j -> (bit, src): skip whole bits, then find the rem-th set
// lane of that bit's ballot by halving.  Requires j < sum of popcounts.
template <int NB>
__device__ inline void kf_job_decode(const unsigned long long bal[NB], int j, int &bit, int &src)
{
    int rem = j;
    unsigned long long m = 0ull;
    bool found = false;
    bit = 0;
#pragma unroll
    for (int b = 0; b < NB; b++) {
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
    src = 0;
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) {
        const int lo = __popcll(m & ((1ull << h) - 1ull));
        if (rem >= lo) {
            rem -= lo;
            m >>= h;
            src += h;
        }
    }
}

inline AdjConsts kf_adj_consts(const kf_params *p)
{
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
    return c;
}

}  // namespace
