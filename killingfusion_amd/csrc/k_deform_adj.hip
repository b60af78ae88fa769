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
// Layout of the launch: a block owns a 64 x 4 (x, y) column of the grid and marches TZ
// slices in z.  The u neighbourhood lives in LDS as three rolling slices (z-1, z, z+1) with
// a one-voxel halo in x and y, so each u value is fetched from memory about once per
// block instead of 13 times per voxel; the live SDF gather is 8 taps straight from L1/L2
// (the sample moves with u).  Blocks are numbered so each XCD sweeps a contiguous z range
// and the slices shared by neighbouring tiles stay in that XCD's L2.
#include "kf_reduce.hpp"

namespace {

constexpr int TX = 64;          // x extent of a tile: one wavefront
constexpr int TY = 4;           // y extent: four wavefronts
constexpr int BLOCK = TX * TY;
constexpr int LW = TX + 2;      // LDS row with x halo
constexpr int LH = TY + 2;      // LDS rows with y halo
constexpr int TZ = 16;          // slices marched per block

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
    // Trilinear value, gradient and mixed second derivatives at base (bx, by, bz) + (fx, fy, fz).
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

__device__ inline float uval(const float *__restrict__ u, const GridDims &g, int c, int x, int y, int z)
{
    return g.in(x, y, z) ? u[(int64_t)c * g.n + g.idx(x, y, z)] : 0.0f;
}

// Fills LDS slot `s` with slice zz of u (zero outside the grid), one-voxel halo in x and y.
__device__ inline void load_slice(float (*su)[3][LH][LW], int s, const float *__restrict__ u, const GridDims &g,
                                  int x0, int y0, int zz, int lane, int wy)
{
    const int x = x0 + lane, y = y0 + wy;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        su[s][c][wy + 1][lane + 1] = uval(u, g, c, x, y, zz);
        if (wy == 0)
            su[s][c][0][lane + 1] = uval(u, g, c, x, y0 - 1, zz);
        if (wy == TY - 1)
            su[s][c][TY + 1][lane + 1] = uval(u, g, c, x, y0 + TY, zz);
    }
    if (lane == 0 || lane == TX - 1) {
        const int col = lane == 0 ? 0 : TX + 1;
        const int xx = lane == 0 ? x0 - 1 : x0 + TX;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            su[s][c][wy + 1][col] = uval(u, g, c, xx, y, zz);
            if (wy == 0)
                su[s][c][0][col] = uval(u, g, c, xx, y0 - 1, zz);
            if (wy == TY - 1)
                su[s][c][TY + 1][col] = uval(u, g, c, xx, y0 + TY, zz);
        }
    }
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void k_adj_step(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                    BlockPartial *partials, float *__restrict__ e_voxel, int ntx,
                                                    int nty)
{
    if (MODE == KF_MODE_STEP && st->done)
        return;
    __shared__ float su[3][3][LH][LW];
    const GridDims g = L.g;
    const unsigned int tile = kf_xcd_tile(blockIdx.x, gridDim.x);
    const int tx = (int)(tile % (unsigned)ntx);
    const int ty = (int)((tile / (unsigned)ntx) % (unsigned)nty);
    const int tz = (int)(tile / ((unsigned)ntx * (unsigned)nty));
    const int lane = threadIdx.x % TX, wy = threadIdx.x / TX;
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int x = x0 + lane, y = y0 + wy;
    const bool valid = x < g.nx && y < g.ny;
    const int zend = min(z0 + TZ, g.nz);

    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;

    int sp = 0, sc = 1, sn = 2;  // LDS slots of slices z-1, z, z+1
    load_slice(su, sp, u_in, g, x0, y0, z0 - 1, lane, wy);
    load_slice(su, sc, u_in, g, x0, y0, z0, lane, wy);
    for (int z = z0; z < zend; z++) {
        load_slice(su, sn, u_in, g, x0, y0, z + 1, lane, wy);
        __syncthreads();
        if (valid) {
            const int64_t i = g.idx(x, y, z);
            const int r = wy + 1, q = lane + 1;
            const float u0 = su[sc][0][r][q], u1 = su[sc][1][r][q], u2 = su[sc][2][r][q];
            // One 8-tap gather gives psi, grad and the mixed second derivatives at j + u_j.
            int bx, by, bz;
            float fx, fy, fz;
            kf_split(x, u0, bx, fx);
            kf_split(y, u1, by, fy);
            kf_split(z, u2, bz, fz);
            float v[8];
            L.cell(bx, by, bz, v);
            const float psi = kf_interp3(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], fx, fy, fz);
            const bool active = !(psi > c.hi || psi < c.lo);
            float gr0 = 0.f, gr1 = 0.f, gr2 = 0.f;
            float e = 0.f;
            if (active) {
                act += 1;
                const float dx00 = v[1] - v[0], dx10 = v[3] - v[2], dx01 = v[5] - v[4], dx11 = v[7] - v[6];
                const float dy00 = v[2] - v[0], dy10 = v[3] - v[1], dy01 = v[6] - v[4], dy11 = v[7] - v[5];
                const float dz00 = v[4] - v[0], dz10 = v[5] - v[1], dz01 = v[6] - v[2], dz11 = v[7] - v[3];
                const float g0 = kf_lerp(kf_lerp(dx00, dx10, fy), kf_lerp(dx01, dx11, fy), fz);
                const float g1 = kf_lerp(kf_lerp(dy00, dy10, fx), kf_lerp(dy01, dy11, fx), fz);
                const float g2 = kf_lerp(kf_lerp(dz00, dz10, fx), kf_lerp(dz01, dz11, fx), fy);
                if (c.use_data) {
                    const float res = psi - Cn.at(x, y, z);
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
                    e += c.omega_s * 0.5f * (sl - 1.0f) * (sl - 1.0f);
                    const float k = c.omega_s * (sl - 1.0f) / sl;
                    gr0 += k * (hxy * g1 + hxz * g2);
                    gr1 += k * (hxy * g0 + hyz * g2);
                    gr2 += k * (hxz * g0 + hyz * g1);
                }
            }
            if (c.use_k) {
                // up[b][a] = u^a_{j+e_b}, um[b][a] = u^a_{j-e_b}; zero outside the grid.
                float up[3][3], um[3][3];
                const float u_c[3] = {u0, u1, u2};
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    up[0][a] = su[sc][a][r][q + 1];
                    um[0][a] = su[sc][a][r][q - 1];
                    up[1][a] = su[sc][a][r + 1][q];
                    um[1][a] = su[sc][a][r - 1][q];
                    up[2][a] = su[sn][a][r][q];
                    um[2][a] = su[sp][a][r][q];
                }
                // dgl[a][b] = u^b_{j+e_a-e_b}, a != b
                float dgl[3][3];
                dgl[0][1] = su[sc][1][r - 1][q + 1];
                dgl[0][2] = su[sp][2][r][q + 1];
                dgl[1][0] = su[sc][0][r + 1][q - 1];
                dgl[1][2] = su[sp][2][r + 1][q];
                dgl[2][0] = su[sn][0][r][q - 1];
                dgl[2][1] = su[sn][1][r - 1][q];
                float kg[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    float lap = 0.f;
#pragma unroll
                    for (int b = 0; b < 3; b++)
                        lap += up[b][a] - 2.0f * u_c[a] + um[b][a];
                    float dv = 0.f;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        // D_a u^b at j and at j - e_b
                        const float d_j = up[a][b] - u_c[b];
                        const float d_m = (a == b) ? (u_c[b] - um[b][b]) : (dgl[a][b] - um[b][b]);
                        dv += d_j - d_m;
                    }
                    kg[a] = -2.0f * (lap + c.gamma * dv);
                }
                gr0 += c.omega_k * kg[0];
                gr1 += c.omega_k * kg[1];
                gr2 += c.omega_k * kg[2];
                if (MODE != KF_MODE_GRAD) {
                    // Energy owned by j: ||D u_j||^2 + gamma tr((D u_j)^2) + ghost terms on low faces.
                    float D[3][3];
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++)
                            D[a][b] = up[b][a] - u_c[a];
                    float ek = 0.f, tr = 0.f;
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) {
                            ek += D[a][b] * D[a][b];
                            tr += D[a][b] * D[b][a];
                        }
                    ek += c.gamma * tr;
                    const int jj[3] = {x, y, z};
                    const float s2 = u0 * u0 + u1 * u1 + u2 * u2;
#pragma unroll
                    for (int cc = 0; cc < 3; cc++)
                        if (jj[cc] == 0)
                            ek += s2 + c.gamma * u_c[cc] * u_c[cc];
                    e += c.omega_k * ek;
                }
            }
            if (MODE == KF_MODE_GRAD) {
                u_out[i] = gr0;
                u_out[g.n + i] = gr1;
                u_out[2 * g.n + i] = gr2;
            } else if (MODE == KF_MODE_ENERGY) {
                if (e_voxel)
                    e_voxel[i] = e;
            } else {
                float n0 = u0, n1 = u1, n2 = u2;
                if (active) {
                    const float d0 = -c.alpha * gr0, d1 = -c.alpha * gr1, d2 = -c.alpha * gr2;
                    n0 = u0 + d0;
                    n1 = u1 + d1;
                    n2 = u2 + d2;
                    if (!(isfinite(n0) && isfinite(n1) && isfinite(n2)))
                        atomicMin((long long *)&st->diverged, (long long)i);
                    else
                        mx = fmax(mx, sqrt((double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2));
                }
                u_out[i] = n0;
                u_out[g.n + i] = n1;
                u_out[2 * g.n + i] = n2;
            }
            esum += (double)e;
        }
        __syncthreads();  // slot sp is refilled next step
        const int t = sp;
        sp = sc;
        sc = sn;
        sn = t;
    }
    if (MODE != KF_MODE_GRAD)
        kf_block_publish<BLOCK>(partials, blockIdx.x, esum, mx, act);
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

}  // namespace

kf_status kf_launch_adj_step(const kf_volume *live, const kf_volume *can, const float *u_in, float *u_out,
                             const kf_params *p, int mode, float *e_voxel)
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
    const int ntx = (live->dims[0] + TX - 1) / TX;
    const int nty = (live->dims[1] + TY - 1) / TY;
    const int ntz = (live->dims[2] + TZ - 1) / TZ;
    const unsigned int nb = (unsigned int)(ntx * nty * ntz);
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        PhiRead<S> L = make_read<S>(live), Cn = make_read<S>(can);
        if (mode == KF_MODE_GRAD)
            hipLaunchKernelGGL((k_adj_step<S, KF_MODE_GRAD>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in,
                               u_out, c, ctx->state, ctx->partials, e_voxel, ntx, nty);
        else if (mode == KF_MODE_ENERGY)
            hipLaunchKernelGGL((k_adj_step<S, KF_MODE_ENERGY>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn,
                               u_in, u_out, c, ctx->state, ctx->partials, e_voxel, ntx, nty);
        else
            hipLaunchKernelGGL((k_adj_step<S, KF_MODE_STEP>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in,
                               u_out, c, ctx->state, ctx->partials, e_voxel, ntx, nty);
    });
    KF_LAUNCH_CHECK();
    return mode == KF_MODE_GRAD ? KF_OK : kf_launch_finalize(ctx, nb);
}
