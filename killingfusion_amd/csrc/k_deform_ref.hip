// k_deform_ref.hip -- the reference energy gradient (KillingFusion::computeEnergyGradient,
// KillingFusion/src/KillingFusion.cpp:399-476) and one all-voxels Jacobi iteration
// (KillingFusion.cpp:248-351) as a gfx950 kernel, FP32 arithmetic.
//
// This is the "plain" kernel: each voxel evaluates the reference's finite-difference
// stencils literally -- 1 data sample, 6 gradient samples and 19 Hessian samples of the
// warped live SDF, each preceded by a trilinear displacement read, plus 19 displacement
// samples for the Killing term -- with every quirk of the reference reads (truncating
// casts, cell-centre offsets, linear-index-only displacement bounds test, clamp before
// the 1/(4h^2) division).
#include "kf_reduce.hpp"

namespace {

constexpr int BLOCK = 256;

template <class S>
struct PhiView {
    const typename S::T *p;
    GridDims g;
    float scale;
    float oob;
    // SDF::getDistanceAtIndex, SDF.cpp:482-487.
    __device__ inline float at(int x, int y, int z) const
    {
        return g.in(x, y, z) ? S::load(p, g.idx(x, y, z), scale) : oob;
    }
    // SDF::getDistance(Vector3d), SDF.cpp:503-524, at the point i + r of the shifted (-0.5)
    // coordinates: base = (int)(i + r), weights = (i + r) - base.
    __device__ inline float trilinear(int ix, int iy, int iz, float rx, float ry, float rz) const
    {
        int bx, by, bz;
        float fx, fy, fz;
        kf_split(ix, rx, bx, fx);
        kf_split(iy, ry, by, fy);
        kf_split(iz, rz, bz, fz);
        return kf_interp3(at(bx, by, bz), at(bx + 1, by, bz), at(bx, by + 1, bz), at(bx + 1, by + 1, bz),
                          at(bx, by, bz + 1), at(bx + 1, by, bz + 1), at(bx, by + 1, bz + 1),
                          at(bx + 1, by + 1, bz + 1), fx, fy, fz);
    }
};

// DisplacementField::getDisplacementAt, DisplacementField.cpp:30-38 (linear-index test only).
__device__ inline void disp_at(const float *__restrict__ u, const GridDims &g, int x, int y, int z, float o[3])
{
    int64_t i = ((int64_t)z * g.ny + y) * g.nx + x;
    if (i < 0 || i >= g.n) {
        o[0] = o[1] = o[2] = 0.0f;
        return;
    }
    o[0] = u[i];
    o[1] = u[g.n + i];
    o[2] = u[2 * g.n + i];
}

// DisplacementField::getDisplacementAtf, DisplacementField.cpp:40-58, at the point i + r.
__device__ inline void disp_at_f(const float *__restrict__ u, const GridDims &g, int ix, int iy, int iz, float rx,
                                 float ry, float rz, float o[3])
{
    int bx, by, bz;
    float fx, fy, fz;
    kf_split(ix, rx, bx, fx);
    kf_split(iy, ry, by, fy);
    kf_split(iz, rz, bz, fz);
    float c[8][3];
    disp_at(u, g, bx, by, bz, c[0]);
    disp_at(u, g, bx + 1, by, bz, c[1]);
    disp_at(u, g, bx, by + 1, bz, c[2]);
    disp_at(u, g, bx + 1, by + 1, bz, c[3]);
    disp_at(u, g, bx, by, bz + 1, c[4]);
    disp_at(u, g, bx + 1, by, bz + 1, c[5]);
    disp_at(u, g, bx, by + 1, bz + 1, c[6]);
    disp_at(u, g, bx + 1, by + 1, bz + 1, c[7]);
#pragma unroll
    for (int k = 0; k < 3; k++)
        o[k] = kf_interp3(c[0][k], c[1][k], c[2][k], c[3][k], c[4][k], c[5][k], c[6][k], c[7][k], fx, fy, fz);
}

// SDF::getDistancef, SDF.cpp:577-582, at loc = i + r: phi at loc + u_f(loc) + 0.5, whose
// shifted coordinate is i + (r + u_f(loc)).
template <class S>
__device__ inline float distance_f(const PhiView<S> &P, const float *__restrict__ u, int ix, int iy, int iz, float rx,
                                   float ry, float rz)
{
    float d[3];
    disp_at_f(u, P.g, ix, iy, iz, rx, ry, rz, d);
    return P.trilinear(ix, iy, iz, rx + d[0], ry + d[1], rz + d[2]);
}

// Offsets (units of h) of the 19-sample stencils, SDF.cpp:847-865 / DisplacementField.cpp:305-323.
__constant__ signed char c_off[19][3] = {
    {0, 0, 0}, {2, 0, 0}, {-2, 0, 0}, {0, 2, 0}, {0, -2, 0}, {0, 0, 2}, {0, 0, -2},
    {1, 0, 1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, 1}, {1, 1, 0}, {-1, 1, 0},
    {-1, 0, -1}, {1, 0, -1}, {0, -1, -1}, {0, 1, -1}, {-1, -1, 0}, {1, -1, 0}};

struct RefConsts {
    float inv_voxel_size_den;  // unused placeholder for alignment
    float voxel_size, hi, lo, eps, alpha, omega_k, gamma, omega_s, h, clamp;
    int use_data, use_ls, use_k;
};

template <class S>
__device__ inline void ref_gradient(const PhiView<S> &L, const PhiView<S> &Cn, const float *__restrict__ u,
                                    const RefConsts &c, int x, int y, int z, float psi, float out[3])
{
    const float h = c.h;
    // Samples sit at gridLocation + offset, gridLocation = idx + 0.5 (SDF.cpp:729,845).
    float dg[3] = {0.f, 0.f, 0.f}, lg[3] = {0.f, 0.f, 0.f}, kg[3] = {0.f, 0.f, 0.f};
    float gr[3] = {0.f, 0.f, 0.f};
    if (c.use_data || c.use_ls) {
        // SDF::computeDistanceGradient(idx, disp), SDF.cpp:726-744.
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float lp[3] = {0.5f, 0.5f, 0.5f}, lm[3] = {0.5f, 0.5f, 0.5f};
            lp[a] = 0.5f + h;
            lm[a] = 0.5f - h;
            gr[a] = distance_f(L, u, x, y, z, lp[0], lp[1], lp[2]) - distance_f(L, u, x, y, z, lm[0], lm[1], lm[2]);
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
            gr[a] = gr[a] / (2.0f * h);
    }
    if (c.use_data) {
        // computeDataEnergyGradient, KillingFusion.cpp:444-455.
        float dst = Cn.at(x, y, z);
        float s = (psi - dst) / c.voxel_size;
#pragma unroll
        for (int a = 0; a < 3; a++)
            dg[a] = s * gr[a];
    }
    if (c.use_ls) {
        // SDF::computeDistanceHessian(idx, disp), SDF.cpp:842-886.
        float f[19];
        for (int k = 0; k < 19; k++)
            f[k] = distance_f(L, u, x, y, z, 0.5f + h * (float)c_off[k][0], 0.5f + h * (float)c_off[k][1],
                              0.5f + h * (float)c_off[k][2]);
        float den = (4.0f * h * h);
        float fxx = (f[1] - 2.0f * f[0] + f[2]) / den;
        float fyy = (f[3] - 2.0f * f[0] + f[4]) / den;
        float fzz = (f[5] - 2.0f * f[0] + f[6]) / den;
        float fxz = (f[7] + f[13] - f[14] - f[8]) / den;
        float fyz = (f[9] + f[15] - f[16] - f[10]) / den;
        float fxy = (f[11] + f[17] - f[18] - f[12]) / den;
        // computeLevelSetEnergyGradient, KillingFusion.cpp:464-476.
        float n = sqrtf(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
        float hg0 = fxx * gr[0] + fxy * gr[1] + fxz * gr[2];
        float hg1 = fxy * gr[0] + fyy * gr[1] + fyz * gr[2];
        float hg2 = fxz * gr[0] + fyz * gr[1] + fzz * gr[2];
        float q = (n - 1.0f);
        float r = (n + c.eps);
        lg[0] = hg0 * q / r * c.omega_s;
        lg[1] = hg1 * q / r * c.omega_s;
        lg[2] = hg2 * q / r * c.omega_s;
    }
    if (c.use_k) {
        // DisplacementField::computeKillingEnergyGradient2, DisplacementField.cpp:301-347.
        float xx[3] = {0, 0, 0}, yy[3] = {0, 0, 0}, zz[3] = {0, 0, 0};
        float xy[3] = {0, 0, 0}, xz[3] = {0, 0, 0}, yz[3] = {0, 0, 0};
        float d0[3];
        disp_at_f(u, L.g, x, y, z, 0.0f, 0.0f, 0.0f, d0);
        float d[3];
        // pure second differences
        disp_at_f(u, L.g, x, y, z, 2.0f * h, 0.0f, 0.0f, d);
        for (int k = 0; k < 3; k++) xx[k] = d[k] - 2.0f * d0[k];
        disp_at_f(u, L.g, x, y, z, -2.0f * h, 0.0f, 0.0f, d);
        for (int k = 0; k < 3; k++) xx[k] = xx[k] + d[k];
        disp_at_f(u, L.g, x, y, z, 0.0f, 2.0f * h, 0.0f, d);
        for (int k = 0; k < 3; k++) yy[k] = d[k] - 2.0f * d0[k];
        disp_at_f(u, L.g, x, y, z, 0.0f, -2.0f * h, 0.0f, d);
        for (int k = 0; k < 3; k++) yy[k] = yy[k] + d[k];
        disp_at_f(u, L.g, x, y, z, 0.0f, 0.0f, 2.0f * h, d);
        for (int k = 0; k < 3; k++) zz[k] = d[k] - 2.0f * d0[k];
        disp_at_f(u, L.g, x, y, z, 0.0f, 0.0f, -2.0f * h, d);
        for (int k = 0; k < 3; k++) zz[k] = zz[k] + d[k];
        // mixed: (+,+) + (-,-) - (+,-) - (-,+)
        float pp[3], mm[3], pm[3], mp[3];
        disp_at_f(u, L.g, x, y, z, h, 0.0f, h, pp);
        disp_at_f(u, L.g, x, y, z, -h, 0.0f, -h, mm);
        disp_at_f(u, L.g, x, y, z, h, 0.0f, -h, pm);
        disp_at_f(u, L.g, x, y, z, -h, 0.0f, h, mp);
        for (int k = 0; k < 3; k++) xz[k] = pp[k] + mm[k] - pm[k] - mp[k];
        disp_at_f(u, L.g, x, y, z, 0.0f, h, h, pp);
        disp_at_f(u, L.g, x, y, z, 0.0f, -h, -h, mm);
        disp_at_f(u, L.g, x, y, z, 0.0f, h, -h, pm);
        disp_at_f(u, L.g, x, y, z, 0.0f, -h, h, mp);
        for (int k = 0; k < 3; k++) yz[k] = pp[k] + mm[k] - pm[k] - mp[k];
        disp_at_f(u, L.g, x, y, z, h, h, 0.0f, pp);
        disp_at_f(u, L.g, x, y, z, -h, -h, 0.0f, mm);
        disp_at_f(u, L.g, x, y, z, h, -h, 0.0f, pm);
        disp_at_f(u, L.g, x, y, z, -h, h, 0.0f, mp);
        for (int k = 0; k < 3; k++) xy[k] = pp[k] + mm[k] - pm[k] - mp[k];
        float div0 = xx[0] + xy[1] + xz[2];
        float div1 = xy[0] + yy[1] + yz[2];
        float div2 = xz[0] + yz[1] + zz[2];
        float g2 = -2.0f * c.gamma;
        kg[0] = -2.0f * (xx[0] + yy[0] + zz[0]) + g2 * div0;
        kg[1] = -2.0f * (xx[1] + yy[1] + zz[1]) + g2 * div1;
        kg[2] = -2.0f * (xx[2] + yy[2] + zz[2]) + g2 * div2;
        float nrm = sqrtf(kg[0] * kg[0] + kg[1] * kg[1] + kg[2] * kg[2]);
        if (nrm > c.clamp) {
            for (int k = 0; k < 3; k++)
                kg[k] = kg[k] / nrm * c.clamp;
        }
        float den = (4.0f * h * h);
        for (int k = 0; k < 3; k++)
            kg[k] = kg[k] / den * c.omega_k;
    }
    // computeEnergyGradient, KillingFusion.cpp:441: data + killing + levelset.
#pragma unroll
    for (int a = 0; a < 3; a++)
        out[a] = (dg[a] + kg[a]) + lg[a];
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void k_ref_step(PhiView<S> L, PhiView<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, RefConsts c, DeformState *st,
                                                    BlockPartial *partials, int parity)
{
    if (MODE == KF_MODE_STEP && st->done)
        return;
    const GridDims g = L.g;
    const int64_t i = (int64_t)kf_xcd_chunk_tile<64>(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    double mx = 0.0;
    unsigned long long act = 0;
    if (i < g.n) {
        int x, y, z;
        kf_xyz(g, i, x, y, z);
        float u0 = u_in[i], u1 = u_in[g.n + i], u2 = u_in[2 * g.n + i];
        // SDF::getDistance(idx, disp), SDF.cpp:570-575; skip test KillingFusion.cpp:272-274.
        float psi = L.trilinear(x, y, z, u0, u1, u2);
        bool active = !(psi > c.hi || psi < c.lo);
        float gr[3] = {0.f, 0.f, 0.f};
        if (active) {
            act = 1;
            ref_gradient<S>(L, Cn, u_in, c, x, y, z, psi, gr);
        }
        if (MODE == KF_MODE_GRAD) {
            u_out[i] = gr[0];
            u_out[g.n + i] = gr[1];
            u_out[2 * g.n + i] = gr[2];
        } else {
            float d0 = -c.alpha * gr[0], d1 = -c.alpha * gr[1], d2 = -c.alpha * gr[2];
            float n0 = u0 + d0, n1 = u1 + d1, n2 = u2 + d2;
            u_out[i] = n0;
            u_out[g.n + i] = n1;
            u_out[2 * g.n + i] = n2;
            if (active) {
                double dn = sqrt((double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2);
                if (!(isfinite(n0) && isfinite(n1) && isfinite(n2)))
                    atomicMin((long long *)&st->div_slot[parity], (long long)i);
                else
                    mx = dn;
            }
        }
    }
    if (MODE == KF_MODE_STEP) {
        double sum = 0.0;
        kf_block_publish<BLOCK>(partials, blockIdx.x, sum, mx, act);
    }
}

template <class S>
PhiView<S> make_view(const kf_volume *v)
{
    PhiView<S> P;
    P.p = (const typename S::T *)v->phi;
    P.g = grid_of(v);
    P.scale = v->scale();
    P.oob = (float)v->oob();
    return P;
}

}  // namespace

kf_status kf_launch_ref_step(const kf_volume *live, const kf_volume *can, const float *u_in, float *u_out,
                             const kf_params *p, int mode, int iter)
{
    kf_context *ctx = live->ctx;
    const int parity = iter & 1;
    RefConsts c{};
    c.voxel_size = (float)p->voxel_size;
    c.hi = (float)(p->max_surface - p->epsilon);
    c.lo = (float)(-p->clip_distance);
    c.eps = (float)p->epsilon;
    c.alpha = (float)p->alpha;
    c.omega_k = (float)p->omega_killing;
    c.gamma = (float)p->gamma_killing;
    c.omega_s = (float)p->omega_levelset;
    c.h = (float)p->delta_size;
    c.clamp = (float)p->killing_clamp;
    c.use_data = p->use_data;
    c.use_ls = p->use_levelset;
    c.use_k = p->use_killing;
    const unsigned int nb = (unsigned int)((live->n + BLOCK - 1) / BLOCK);
    kf_status s = kf_ensure_partials(ctx, nb);
    if (s != KF_OK)
        return s;
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    KF_DISPATCH_DTYPE(live->dtype, S, {
        PhiView<S> L = make_view<S>(live), Cn = make_view<S>(can);
        if (mode == KF_MODE_GRAD)
            hipLaunchKernelGGL((k_ref_step<S, KF_MODE_GRAD>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in,
                               u_out, c, ctx->state, kf_partials(ctx, parity), parity);
        else
            hipLaunchKernelGGL((k_ref_step<S, KF_MODE_STEP>), dim3(nb), dim3(BLOCK), 0, ctx->stream, L, Cn, u_in,
                               u_out, c, ctx->state, kf_partials(ctx, parity), parity);
    });
    KF_LAUNCH_CHECK();
    return mode == KF_MODE_GRAD ? KF_OK : kf_launch_finalize(ctx, nb, parity);
}
