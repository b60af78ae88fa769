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
#include "kf_reduce.hpp"

namespace {

constexpr int BLOCK = 256;

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
};

__device__ inline float uval(const float *__restrict__ u, const GridDims &g, int c, int x, int y, int z)
{
    return g.in(x, y, z) ? u[(int64_t)c * g.n + g.idx(x, y, z)] : 0.0f;
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void k_adj_step(PhiRead<S> L, PhiRead<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, AdjConsts c, DeformState *st,
                                                    BlockPartial *partials, float *__restrict__ e_voxel)
{
    if (MODE == KF_MODE_STEP && st->done)
        return;
    const GridDims g = L.g;
    const int64_t i = (int64_t)kf_xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    double mx = 0.0, esum = 0.0;
    unsigned long long act = 0;
    if (i < g.n) {
        const int x = (int)(i % g.nx);
        const int y = (int)((i / g.nx) % g.ny);
        const int z = (int)(i / ((int64_t)g.nx * g.ny));
        const float u0 = u_in[i], u1 = u_in[g.n + i], u2 = u_in[2 * g.n + i];
        // One 8-tap gather gives psi, grad and the mixed second derivatives.
        int bx, by, bz;
        float fx, fy, fz;
        kf_split(x, u0, bx, fx);
        kf_split(y, u1, by, fy);
        kf_split(z, u2, bz, fz);
        const float v000 = L.at(bx, by, bz), v001 = L.at(bx + 1, by, bz);
        const float v010 = L.at(bx, by + 1, bz), v011 = L.at(bx + 1, by + 1, bz);
        const float v100 = L.at(bx, by, bz + 1), v101 = L.at(bx + 1, by, bz + 1);
        const float v110 = L.at(bx, by + 1, bz + 1), v111 = L.at(bx + 1, by + 1, bz + 1);
        const float psi = kf_interp3(v000, v001, v010, v011, v100, v101, v110, v111, fx, fy, fz);
        const bool active = !(psi > c.hi || psi < c.lo);
        float gr0 = 0.f, gr1 = 0.f, gr2 = 0.f;
        float e = 0.f;
        if (active) {
            act = 1;
            const float dx00 = v001 - v000, dx10 = v011 - v010, dx01 = v101 - v100, dx11 = v111 - v110;
            const float dy00 = v010 - v000, dy10 = v011 - v001, dy01 = v110 - v100, dy11 = v111 - v101;
            const float dz00 = v100 - v000, dz10 = v101 - v001, dz01 = v110 - v010, dz11 = v111 - v011;
            const float g0 = kf_lerp(kf_lerp(dx00, dx10, fy), kf_lerp(dx01, dx11, fy), fz);
            const float g1 = kf_lerp(kf_lerp(dy00, dy10, fx), kf_lerp(dy01, dy11, fx), fz);
            const float g2 = kf_lerp(kf_lerp(dz00, dz10, fx), kf_lerp(dz01, dz11, fx), fy);
            if (c.use_data) {
                const float r = psi - Cn.at(x, y, z);
                e += r * r * c.inv_2v;
                const float s = r * c.inv_v;
                gr0 += s * g0;
                gr1 += s * g1;
                gr2 += s * g2;
            }
            if (c.use_ls) {
                const float hxy = kf_lerp(dx10 - dx00, dx11 - dx01, fz);
                const float hxz = kf_lerp(dx01 - dx00, dx11 - dx10, fy);
                const float hyz = kf_lerp(dy01 - dy00, dy11 - dy10, fx);
                const float sn = sqrtf(g0 * g0 + g1 * g1 + g2 * g2 + c.eps2);
                e += c.omega_s * 0.5f * (sn - 1.0f) * (sn - 1.0f);
                const float k = c.omega_s * (sn - 1.0f) / sn;
                gr0 += k * (hxy * g1 + hxz * g2);
                gr1 += k * (hxy * g0 + hyz * g2);
                gr2 += k * (hxz * g0 + hyz * g1);
            }
        }
        if (c.use_k) {
            // Neighbourhood of u (zero outside the grid).
            float up[3][3], um[3][3];  // up[b][a] = u^a_{j+e_b}, um[b][a] = u^a_{j-e_b}
            const float u_c[3] = {u0, u1, u2};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                up[0][a] = uval(u_in, g, a, x + 1, y, z);
                um[0][a] = uval(u_in, g, a, x - 1, y, z);
                up[1][a] = uval(u_in, g, a, x, y + 1, z);
                um[1][a] = uval(u_in, g, a, x, y - 1, z);
                up[2][a] = uval(u_in, g, a, x, y, z + 1);
                um[2][a] = uval(u_in, g, a, x, y, z - 1);
            }
            // diag[a][b] = u^b_{j+e_a-e_b}, a != b
            float dgl[3][3];
            dgl[0][1] = uval(u_in, g, 1, x + 1, y - 1, z);
            dgl[0][2] = uval(u_in, g, 2, x + 1, y, z - 1);
            dgl[1][0] = uval(u_in, g, 0, x - 1, y + 1, z);
            dgl[1][2] = uval(u_in, g, 2, x, y + 1, z - 1);
            dgl[2][0] = uval(u_in, g, 0, x - 1, y, z + 1);
            dgl[2][1] = uval(u_in, g, 1, x, y - 1, z + 1);
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
            act = active ? 1 : 0;
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
                    mx = sqrt((double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2);
            }
            u_out[i] = n0;
            u_out[g.n + i] = n1;
            u_out[2 * g.n + i] = n2;
        }
        esum = (double)e;
    }
    if (MODE != KF_MODE_GRAD) {
        kf_block_publish<BLOCK>(partials, blockIdx.x, esum, mx, act);
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
    const unsigned int nb = (unsigned int)((live->n + BLOCK - 1) / BLOCK);
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    KF_DISPATCH_DTYPE(live->dtype, S, {
        PhiRead<S> L = make_read<S>(live), Cn = make_read<S>(can);
        if (mode == KF_MODE_GRAD)
            hipLaunchKernelGGL((k_adj_step<S, KF_MODE_GRAD>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in,
                               u_out, c, ctx->state, ctx->partials, e_voxel);
        else if (mode == KF_MODE_ENERGY)
            hipLaunchKernelGGL((k_adj_step<S, KF_MODE_ENERGY>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn,
                               u_in, u_out, c, ctx->state, ctx->partials, e_voxel);
        else
            hipLaunchKernelGGL((k_adj_step<S, KF_MODE_STEP>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in,
                               u_out, c, ctx->state, ctx->partials, e_voxel);
    });
    KF_LAUNCH_CHECK();
    return mode == KF_MODE_GRAD ? KF_OK : kf_launch_finalize(ctx, nb);
}
