// k_deform_ref.hip -- the reference energy gradient (KillingFusion::computeEnergyGradient,
// KillingFusion/src/KillingFusion.cpp:399-476) and one all-voxels Jacobi iteration
// (KillingFusion.cpp:248-351) on gfx950, FP32 arithmetic.
//
// This is the "plain" path: each voxel evaluates the reference's finite-difference
// stencils literally -- 1 data sample, 6 gradient samples and 19 Hessian samples of the
// warped live SDF, each preceded by a trilinear displacement read, plus 19 displacement
// samples for the Killing term -- with every quirk of the reference reads (truncating
// casts, cell-centre offsets, linear-index-only displacement bounds test, clamp before
// the 1/(4h^2) division).  A step runs in two passes: k_ref_scan streams every voxel, tests
// the band and lists the voxels in it; k_ref_list takes the listed voxels' samples, one
// (voxel, sample) pair per thread, and combines them.  The results do not depend on this
// layout: each sample and each combination is the same per-voxel arithmetic.
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
        float v[8];
        if (bx >= 0 && by >= 0 && bz >= 0 && bx + 1 < g.nx && by + 1 < g.ny && bz + 1 < g.nz) {
            // the whole cell in the grid: one bounds test, taps at fixed strides
            const int64_t i0 = g.idx(bx, by, bz), sy = g.nx, sz = (int64_t)g.nx * g.ny;
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
        return kf_interp3(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], fx, fy, fz);
    }
    // SDF::getDistance(Vector3d), SDF.cpp:503-524, in the reference's own FP64 arithmetic at grid
    // location loc: loc - 0.5, truncating casts, the lerps in double, taps past the grid reading
    // oob64 (Max + eps in FP64).  A non-finite location reads NaN; one beyond 2^30 voxels reads
    // only taps past the grid (the reference's cast is undefined there).
    __device__ inline double distance64(const double loc[3], double oob64) const
    {
        double t[3];
        int b[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            t[a] = loc[a] - 0.5;
            if (!isfinite(t[a]))
                return (double)NAN;
            const double c = fmin(fmax(t[a], -1073741824.0), 1073741824.0);
            b[a] = (int)c;
        }
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int x = b[0] + (k & 1), y = b[1] + ((k >> 1) & 1), z = b[2] + (k >> 2);
            v[k] = g.in(x, y, z) ? (double)S::load(p, g.idx(x, y, z), scale) : oob64;
        }
        return kf_interp3(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], t[0] - (double)b[0], t[1] - (double)b[1],
                          t[2] - (double)b[2]);
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
// NB: the taps come from nb, the 3x3x3 neighbourhood of (ix, iy, iz) read by disp_at (entry
// ((dz + 1) * 3 + dy + 1) * 3 + dx + 1 holds u at (ix + dx, iy + dy, iz + dz)); the caller
// guarantees |r| < 1 on each axis, so the cell lies in it.
template <bool NB = false>
__device__ inline void disp_at_f(const float *__restrict__ u, const GridDims &g, int ix, int iy, int iz, float rx,
                                 float ry, float rz, float o[3], const float (*nb)[3] = nullptr)
{
    int bx, by, bz;
    float fx, fy, fz;
    kf_split(ix, rx, bx, fx);
    kf_split(iy, ry, by, fy);
    kf_split(iz, rz, bz, fz);
    float c[8][3];
    if constexpr (NB) {
        const int e = ((bz - iz + 1) * 3 + (by - iy + 1)) * 3 + (bx - ix + 1);
        constexpr int off[8] = {0, 1, 3, 4, 9, 10, 12, 13};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            c[k][0] = nb[e + off[k]][0];
            c[k][1] = nb[e + off[k]][1];
            c[k][2] = nb[e + off[k]][2];
        }
    } else {
        disp_at(u, g, bx, by, bz, c[0]);
        disp_at(u, g, bx + 1, by, bz, c[1]);
        disp_at(u, g, bx, by + 1, bz, c[2]);
        disp_at(u, g, bx + 1, by + 1, bz, c[3]);
        disp_at(u, g, bx, by, bz + 1, c[4]);
        disp_at(u, g, bx + 1, by, bz + 1, c[5]);
        disp_at(u, g, bx, by + 1, bz + 1, c[6]);
        disp_at(u, g, bx + 1, by + 1, bz + 1, c[7]);
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        o[k] = kf_interp3(c[0][k], c[1][k], c[2][k], c[3][k], c[4][k], c[5][k], c[6][k], c[7][k], fx, fy, fz);
}

// Offsets (units of h) of the 19-sample Hessian stencil, SDF.cpp:847-865.
__constant__ signed char c_off[19][3] = {
    {0, 0, 0}, {2, 0, 0}, {-2, 0, 0}, {0, 2, 0}, {0, -2, 0}, {0, 0, 2}, {0, 0, -2},
    {1, 0, 1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, 1}, {1, 1, 0}, {-1, 1, 0},
    {-1, 0, -1}, {1, 0, -1}, {0, -1, -1}, {0, 1, -1}, {-1, -1, 0}, {1, -1, 0}};

// Offsets (units of h) of the 19 displacement samples of the Killing gradient, in the order
// DisplacementField.cpp:305-323 combines them: the centre, the pure second differences
// (+2h, -2h per axis), then (+,+), (-,-), (+,-), (-,+) of the xz, yz and xy planes.
__constant__ signed char c_koff[19][3] = {
    {0, 0, 0}, {2, 0, 0}, {-2, 0, 0}, {0, 2, 0}, {0, -2, 0}, {0, 0, 2}, {0, 0, -2},
    {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};

// Samples an active voxel takes.  Samples 0-5 are the warped live SDF at the centre +h / -h
// along x, y, z (SDF::computeDistanceGradient, SDF.cpp:726-744), 6-24 the Hessian stencil
// (SDF.cpp:842-886) and 25-43 the displacement at the Killing stencil points.
constexpr int NPHI = 25;
constexpr int NKIL = 19;
// Active voxels whose samples one round of a list block takes, and the list block's threads.
constexpr int RK = 64;
constexpr int LBLOCK = 256;

struct RefConsts {
    float voxel_size, hi, lo, eps, alpha, omega_k, gamma, omega_s, h, clamp;
    int use_data, use_ls, use_k;
    int nphi;   // SDF samples in use: 0, 6 (gradient) or 25 (gradient and Hessian)
    int nsamp;  // nphi + 19 when the Killing term is in use
    int trust;  // the trust region (UseTrustStrategy with the data energy alone)
    double alpha64;  // alpha as the trust region shrinks it (FP64, as the reference's _alpha)
    double oob64;    // the live volume's Max + eps in FP64, for the trust region's reads
};

// Sample s of the active voxel (x, y, z) at the point i + r.  Every sample reads u
// trilinearly there (DisplacementField::getDisplacementAtf); an SDF sample then reads the live
// SDF at the point moved by it (SDF::getDistancef, SDF.cpp:577-582: phi at loc + u_f(loc) +
// 0.5, whose shifted coordinate is i + (r + u_f)).
template <class S, bool NB>
__device__ inline void ref_sample(const PhiView<S> &L, const float *__restrict__ u, float h, int s, int x, int y,
                                  int z, float *phi_out, float *u_out, const float (*nb)[3])
{
    float r0, r1, r2;
    if (s < 6) {
        const int a = s >> 1;
        const float d = (s & 1) ? -h : h;
        r0 = a == 0 ? 0.5f + d : 0.5f;
        r1 = a == 1 ? 0.5f + d : 0.5f;
        r2 = a == 2 ? 0.5f + d : 0.5f;
    } else if (s < NPHI) {
        const int k = s - 6;
        r0 = 0.5f + h * (float)c_off[k][0];
        r1 = 0.5f + h * (float)c_off[k][1];
        r2 = 0.5f + h * (float)c_off[k][2];
    } else {
        const int k = s - NPHI;
        r0 = h * (float)c_koff[k][0];
        r1 = h * (float)c_koff[k][1];
        r2 = h * (float)c_koff[k][2];
    }
    float d[3];
    disp_at_f<NB>(u, L.g, x, y, z, r0, r1, r2, d, nb);
    if (s < NPHI) {
        phi_out[s] = L.trilinear(x, y, z, r0 + d[0], r1 + d[1], r2 + d[2]);
    } else {
        float *o = u_out + 3 * (s - NPHI);
        o[0] = d[0];
        o[1] = d[1];
        o[2] = d[2];
    }
}

// KillingFusion::computeEnergyGradient (KillingFusion.cpp:399-442) from the samples of one
// voxel: f[0..24] the SDF samples, k[0..18] the Killing displacement samples.
__device__ inline void ref_combine(const RefConsts &c, float psi, float dst, const float *f, const float (*k)[3],
                                   float out[3])
{
    const float h = c.h;
    float dg[3] = {0.f, 0.f, 0.f}, lg[3] = {0.f, 0.f, 0.f}, kg[3] = {0.f, 0.f, 0.f};
    float gr[3] = {0.f, 0.f, 0.f};
    if (c.use_data || c.use_ls) {
        // SDF::computeDistanceGradient(idx, disp), SDF.cpp:726-744.
#pragma unroll
        for (int a = 0; a < 3; a++)
            gr[a] = f[2 * a] - f[2 * a + 1];
#pragma unroll
        for (int a = 0; a < 3; a++)
            gr[a] = gr[a] / (2.0f * h);
    }
    if (c.use_data) {
        // computeDataEnergyGradient, KillingFusion.cpp:444-455.
        float s = (psi - dst) / c.voxel_size;
#pragma unroll
        for (int a = 0; a < 3; a++)
            dg[a] = s * gr[a];
    }
    if (c.use_ls) {
        // SDF::computeDistanceHessian(idx, disp), SDF.cpp:842-886.
        const float *q = f + 6;
        float den = (4.0f * h * h);
        float fxx = (q[1] - 2.0f * q[0] + q[2]) / den;
        float fyy = (q[3] - 2.0f * q[0] + q[4]) / den;
        float fzz = (q[5] - 2.0f * q[0] + q[6]) / den;
        float fxz = (q[7] + q[13] - q[14] - q[8]) / den;
        float fyz = (q[9] + q[15] - q[16] - q[10]) / den;
        float fxy = (q[11] + q[17] - q[18] - q[12]) / den;
        // computeLevelSetEnergyGradient, KillingFusion.cpp:464-476.
        float n = sqrtf(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
        float hg0 = fxx * gr[0] + fxy * gr[1] + fxz * gr[2];
        float hg1 = fxy * gr[0] + fyy * gr[1] + fyz * gr[2];
        float hg2 = fxz * gr[0] + fyz * gr[1] + fzz * gr[2];
        float qn = (n - 1.0f);
        float r = (n + c.eps);
        lg[0] = hg0 * qn / r * c.omega_s;
        lg[1] = hg1 * qn / r * c.omega_s;
        lg[2] = hg2 * qn / r * c.omega_s;
    }
    if (c.use_k) {
        // DisplacementField::computeKillingEnergyGradient2, DisplacementField.cpp:301-347.
        float xx[3], yy[3], zz[3], xy[3], xz[3], yz[3];
        const float *d0 = k[0];
        for (int j = 0; j < 3; j++) {
            // pure second differences
            xx[j] = k[1][j] - 2.0f * d0[j];
            xx[j] = xx[j] + k[2][j];
            yy[j] = k[3][j] - 2.0f * d0[j];
            yy[j] = yy[j] + k[4][j];
            zz[j] = k[5][j] - 2.0f * d0[j];
            zz[j] = zz[j] + k[6][j];
            // mixed: (+,+) + (-,-) - (+,-) - (-,+)
            xz[j] = k[7][j] + k[8][j] - k[9][j] - k[10][j];
            yz[j] = k[11][j] + k[12][j] - k[13][j] - k[14][j];
            xy[j] = k[15][j] + k[16][j] - k[17][j] - k[18][j];
        }
        float div0 = xx[0] + xy[1] + xz[2];
        float div1 = xy[0] + yy[1] + yz[2];
        float div2 = xz[0] + yz[1] + zz[2];
        float g2 = -2.0f * c.gamma;
        kg[0] = -2.0f * (xx[0] + yy[0] + zz[0]) + g2 * div0;
        kg[1] = -2.0f * (xx[1] + yy[1] + zz[1]) + g2 * div1;
        kg[2] = -2.0f * (xx[2] + yy[2] + zz[2]) + g2 * div2;
        float nrm = sqrtf(kg[0] * kg[0] + kg[1] * kg[1] + kg[2] * kg[2]);
        if (nrm > c.clamp) {
            for (int j = 0; j < 3; j++)
                kg[j] = kg[j] / nrm * c.clamp;
        }
        float den = (4.0f * h * h);
        for (int j = 0; j < 3; j++)
            kg[j] = kg[j] / den * c.omega_k;
    }
    // computeEnergyGradient, KillingFusion.cpp:441: data + killing + levelset.
#pragma unroll
    for (int a = 0; a < 3; a++)
        out[a] = (dg[a] + kg[a]) + lg[a];
}

// Threads of a pass-1 block.
constexpr int SBLOCK = 256;
// Blocks of the list kernel (at most; fewer for small grids).
constexpr int LIST_BLOCKS = 2048;
// The list has one segment per XCD (block b of pass 1 runs on XCD b % 8), each with its own
// length counter, so the appends of the eight XCDs do not meet on one address.  Workspace
// words: the segment lengths of launch parity 0, those of parity 1, the non-finite flag.
constexpr int NSEG = 8;
constexpr int AUX_WORDS = 2 * NSEG + 1;
constexpr int AUX_NONFINITE = 2 * NSEG;

__device__ inline float4 ref_load_stream4(const float *p)
{
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// The same for a group that starts on a 4-byte boundary only (one 16-byte access on gfx950): the
// u planes a N (a = 1, 2) when N is not a multiple of 4.
typedef float ref_f4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ inline float4 ref_load_stream4u(const float *p)
{
    const ref_f4u v = __builtin_nontemporal_load(reinterpret_cast<const ref_f4u *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Clears the list lengths and the non-finite flag of the workspace.
__global__ void k_ref_clear(unsigned long long *aux)
{
    if (threadIdx.x < AUX_WORDS)
        aux[threadIdx.x] = 0ull;
}

// Slot of list entry p: segments hold pre[x+1] - pre[x] entries from x * segcap on.
__device__ inline unsigned long long ref_slot(const unsigned long long pre[NSEG + 1], unsigned long long segcap,
                                              unsigned long long p)
{
    int x = 0;
#pragma unroll
    for (int k = 1; k < NSEG; k++)
        x += p >= pre[k] ? 1 : 0;
    return (unsigned long long)x * segcap + (p - pre[x]);
}

// Sets *flag when some live SDF value, decoded, is not finite.  V values per thread and step
// (4: one vector load; the grid's voxel count is then a multiple of 4).
template <class S, int V>
__global__ __launch_bounds__(BLOCK) void k_ref_finite(PhiView<S> L, unsigned long long *flag)
{
    bool bad = false;
    const int64_t stride = (int64_t)gridDim.x * BLOCK * V;
    for (int64_t i = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * V; i < L.g.n; i += stride) {
        float v[V];
        if constexpr (V == 4)
            S::load4(L.p, i, L.scale, v);
        else
            v[0] = S::load(L.p, i, L.scale);
#pragma unroll
        for (int k = 0; k < V; k++)
            bad |= !isfinite(v[k]);
    }
    if (__ballot(bad) != 0ull && __lane_id() == 0)
        atomicOr(flag, 1ull);
}

// Pass 1 of a reference-formula step: every voxel's warped live SDF value
// psi = SDF::getDistance(idx, disp) (SDF.cpp:570-575) and the band test
// (KillingFusion.cpp:272-274), 4 voxels per thread (one float4 of each u plane; UNAL: N is not a
// multiple of 4, so planes 1 and 2 are read as 4-byte-aligned 16-byte accesses and the group
// that runs past N goes element by element).  A voxel
// outside the band takes no update: u plus the +0 its iteration's fresh delta field holds
// (KillingFusion.cpp:254-256, 345-348), so a -0 component becomes +0.  A voxel in the band is appended
// to its XCD's list segment for pass 2 (one atomic per block), which overwrites its entry of
// u_out; the list's length is the launch's active count.
// Where u_j is +0 or -0 the interpolation weights are +0 and the interpolant is
// phi_j + (finite) * 0 + ...; when every live value is finite (no non-finite flag) that is phi_j
// itself unless phi_j is -0, so those voxels read phi_j alone.
template <class S, int MODE, bool UNAL>
__global__ __launch_bounds__(SBLOCK) void k_ref_scan(PhiView<S> L, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, float hi, float lo, float alpha,
                                                    const DeformState *st, unsigned long long *aux, int parity,
                                                    unsigned long long segcap,
                                                    unsigned long long *__restrict__ list_idx,
                                                    float *__restrict__ list_psi)
{
    if (MODE == KF_MODE_STEP && st->done)
        return;
    __shared__ unsigned int s_wtot[SBLOCK / 64];
    __shared__ unsigned long long s_base;
    const GridDims g = L.g;
    constexpr int V = 4;
    const int64_t i0 = ((int64_t)kf_xcd_chunk_tile<64>(blockIdx.x, gridDim.x) * SBLOCK + threadIdx.x) * V;
    const bool phi_finite = aux[AUX_NONFINITE] == 0ull;
    bool act[V];
    float psi[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        act[v] = false;
        psi[v] = 0.0f;
    }
    // voxels of the group inside the grid (4 unless UNAL and the group runs past N)
    const int nin = i0 >= g.n ? 0 : (UNAL && g.n - i0 < V ? (int)(g.n - i0) : V);
    float u[3][V];
    if (nin > 0) {
        float ph[V];
        if (nin == V) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float4 q = UNAL ? ref_load_stream4u(u_in + a * g.n + i0) : ref_load_stream4(u_in + a * g.n + i0);
                u[a][0] = q.x;
                u[a][1] = q.y;
                u[a][2] = q.z;
                u[a][3] = q.w;
            }
            S::load4(L.p, i0, L.scale, ph);
        } else {
#pragma unroll
            for (int v = 0; v < V; v++) {
#pragma unroll
                for (int a = 0; a < 3; a++)
                    u[a][v] = v < nin ? u_in[a * g.n + i0 + v] : 0.0f;
                ph[v] = v < nin ? S::load(L.p, i0 + v, L.scale) : 0.0f;
            }
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (v >= nin)
                break;
            const unsigned int zb =
                (__float_as_uint(u[0][v]) | __float_as_uint(u[1][v]) | __float_as_uint(u[2][v])) & 0x7fffffffu;
            if (phi_finite && zb == 0u && __float_as_uint(ph[v]) != 0x80000000u) {
                psi[v] = ph[v];
            } else {
                int x, y, z;
                kf_xyz(g, i0 + v, x, y, z);
                psi[v] = L.trilinear(x, y, z, u[0][v], u[1][v], u[2][v]);
            }
            act[v] = !(psi[v] > hi || psi[v] < lo);
        }
    }
    unsigned long long m[V];
    unsigned int tot = 0;
#pragma unroll
    for (int v = 0; v < V; v++) {
        m[v] = __ballot(act[v]);
        tot += (unsigned int)__popcll(m[v]);
    }
    const int wv = (int)(threadIdx.x >> 6);
    if (__lane_id() == 0)
        s_wtot[wv] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int bt = 0;
#pragma unroll
        for (int k = 0; k < SBLOCK / 64; k++)
            bt += s_wtot[k];
        const unsigned int seg = blockIdx.x % NSEG;
        s_base = bt ? seg * segcap + atomicAdd(&aux[parity * NSEG + seg], (unsigned long long)bt) : 0ull;
    }
    __syncthreads();
    if (tot != 0u) {
        unsigned long long pos = s_base;
        for (int k = 0; k < wv; k++)
            pos += s_wtot[k];
        const unsigned long long below = (1ull << __lane_id()) - 1ull;
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (act[v]) {
                const unsigned long long q = pos + (unsigned long long)__popcll(m[v] & below);
                list_idx[q] = (unsigned long long)(i0 + v);
                list_psi[q] = psi[v];
            }
            pos += (unsigned long long)__popcll(m[v]);
        }
    }
    // The field values go out last: the wait on the list counter above would otherwise also
    // wait for these stores (one counter tracks both on this target) while the block's other
    // waves sit at the barrier.  Voxels in the band get their value in pass 2.
    if (nin > 0) {
        const float d = 0.0f;  // the fresh delta field's +0: a -0 component becomes +0
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float o[V];
#pragma unroll
            for (int v = 0; v < V; v++)
                o[v] = MODE == KF_MODE_GRAD ? 0.0f : u[a][v] + d;
            float *dst = u_out + a * g.n + i0;
            if (!UNAL) {
                *reinterpret_cast<float4 *>(dst) = make_float4(o[0], o[1], o[2], o[3]);
            } else if (nin == V) {
                *reinterpret_cast<ref_f4u *>(dst) = ref_f4u{o[0], o[1], o[2], o[3]};
            } else {
#pragma unroll
                for (int v = 0; v < V; v++)
                    if (v < nin)
                        dst[v] = o[v];
            }
        }
    }
}

// Pass 1 over the voxels that can move only (the active-set sweep; vlist holds them, st->list_n
// of them, within the launch's slots): the same per-voxel work as k_ref_scan, one voxel per
// thread.  Every other voxel holds
// u = +0 in both Jacobi buffers and keeps it (its band test reads phi_live[j] alone and fails,
// or its warped value is NaN), so neither buffer needs a store there.
template <class S, int MODE>
__global__ __launch_bounds__(SBLOCK) void k_ref_active_scan(PhiView<S> L, const float *__restrict__ u_in,
                                                           float *__restrict__ u_out, float hi, float lo,
                                                           float alpha, const DeformState *st,
                                                           unsigned long long *aux, int parity,
                                                           unsigned long long segcap,
                                                           unsigned long long *__restrict__ list_idx,
                                                           float *__restrict__ list_psi,
                                                           const unsigned int *__restrict__ vlist)
{
    const unsigned int count = st->list_n;
    if (MODE == KF_MODE_STEP && st->done)
        return;
    __shared__ unsigned int s_wtot[SBLOCK / 64];
    __shared__ unsigned long long s_base;
    const GridDims g = L.g;
    const unsigned int t = blockIdx.x * SBLOCK + threadIdx.x;
    const bool have = t < count;
    const int64_t j = have ? (int64_t)vlist[t] : 0;
    const bool phi_finite = aux[AUX_NONFINITE] == 0ull;
    float u[3] = {0.0f, 0.0f, 0.0f};
    float psi = 0.0f;
    bool act = false;
    if (have) {
#pragma unroll
        for (int a = 0; a < 3; a++)
            u[a] = u_in[a * g.n + j];
        const float ph = S::load(L.p, j, L.scale);
        const unsigned int zb = (__float_as_uint(u[0]) | __float_as_uint(u[1]) | __float_as_uint(u[2])) & 0x7fffffffu;
        if (phi_finite && zb == 0u && __float_as_uint(ph) != 0x80000000u) {
            psi = ph;
        } else {
            int x, y, z;
            kf_xyz(g, j, x, y, z);
            psi = L.trilinear(x, y, z, u[0], u[1], u[2]);
        }
        act = !(psi > hi || psi < lo);
    }
    const unsigned long long m = __ballot(act);
    const unsigned int tot = (unsigned int)__popcll(m);
    const int wv = (int)(threadIdx.x >> 6);
    if (__lane_id() == 0)
        s_wtot[wv] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int bt = 0;
#pragma unroll
        for (int k = 0; k < SBLOCK / 64; k++)
            bt += s_wtot[k];
        const unsigned int seg = blockIdx.x % NSEG;
        s_base = bt ? seg * segcap + atomicAdd(&aux[parity * NSEG + seg], (unsigned long long)bt) : 0ull;
    }
    __syncthreads();
    if (act) {
        unsigned long long pos = s_base;
        for (int k = 0; k < wv; k++)
            pos += s_wtot[k];
        const unsigned long long q = pos + (unsigned long long)__popcll(m & ((1ull << __lane_id()) - 1ull));
        list_idx[q] = (unsigned long long)j;
        list_psi[q] = psi;
    }
    // field values last (see k_ref_scan); voxels in the band get theirs in pass 2
    if (have) {
        const float d = 0.0f;  // the fresh delta field's +0: a -0 component becomes +0
#pragma unroll
        for (int a = 0; a < 3; a++)
            u_out[a * g.n + j] = MODE == KF_MODE_GRAD ? 0.0f : u[a] + d;
    }
}

// Pass 2: the voxels of the list, RK at a time per block.  The block takes their samples, one
// (voxel, sample) pair per thread, so a voxel's 44 dependent reads run side by side instead of
// one after another in one lane, and a thread per voxel combines them.  Each sample and the
// combination are the per-voxel arithmetic of the reference's stencils, so the result does not
// depend on the list's order or on how the pairs are spread.  A step launch publishes the
// list's length as the active count (block 0, thread 0) and clears the other list length, which the next
// launch's pass 1 counts into.
// NB (0 < h < 1/4): every sample's u cell lies in the voxel's 3x3x3 neighbourhood (the SDF
// samples sit at i + 0.5 +- up to 2h, the Killing samples at i +- up to 2h), so the block reads
// that neighbourhood once per voxel, under the same linear-index test, and the samples take
// their u taps from LDS: 81 loads per voxel instead of 24 per sample.
template <class S, int MODE, bool NB>
__global__ __launch_bounds__(LBLOCK) void k_ref_list(PhiView<S> L, PhiView<S> Cn, const float *__restrict__ u_in,
                                                    float *__restrict__ u_out, RefConsts c, DeformState *st,
                                                    BlockPartial *partials, unsigned long long *aux, int parity,
                                                    unsigned long long segcap,
                                                    const unsigned long long *__restrict__ list_idx,
                                                    float *__restrict__ list_psi, unsigned int *__restrict__ row_fail)
{
    if (MODE == KF_MODE_STEP && st->done)
        return;
    __shared__ long long s_i[RK];
    __shared__ unsigned long long s_q[RK];
    __shared__ int s_xyz[RK][3];
    __shared__ float s_u[RK][3];
    __shared__ float s_psi[RK], s_dst[RK];
    __shared__ float s_f[RK][NPHI];
    __shared__ float s_k[RK][NKIL][3];
    __shared__ float s_nb[NB ? RK : 1][27][3];
    const GridDims g = L.g;
    unsigned long long pre[NSEG + 1];
    pre[0] = 0ull;
#pragma unroll
    for (int k = 0; k < NSEG; k++)
        pre[k + 1] = pre[k] + aux[parity * NSEG + k];
    const unsigned long long count = pre[NSEG];
    if (MODE == KF_MODE_STEP && blockIdx.x == 0 && threadIdx.x < NSEG)
        aux[(parity ^ 1) * NSEG + threadIdx.x] = 0ull;
    double mx = 0.0;
    for (unsigned long long c0 = (unsigned long long)blockIdx.x * RK; c0 < count;
         c0 += (unsigned long long)gridDim.x * RK) {
        const int nv = (int)(count - c0 < (unsigned long long)RK ? count - c0 : (unsigned long long)RK);
        if ((int)threadIdx.x < nv) {
            const int t = threadIdx.x;
            const unsigned long long q = ref_slot(pre, segcap, c0 + t);
            const long long j = (long long)list_idx[q];
            int x, y, z;
            kf_xyz(g, j, x, y, z);
            s_i[t] = j;
            s_q[t] = q;
            s_xyz[t][0] = x;
            s_xyz[t][1] = y;
            s_xyz[t][2] = z;
            s_u[t][0] = u_in[j];
            s_u[t][1] = u_in[g.n + j];
            s_u[t][2] = u_in[2 * g.n + j];
            s_psi[t] = list_psi[q];
            s_dst[t] = c.use_data ? Cn.at(x, y, z) : 0.0f;
        }
        if constexpr (NB) {
#pragma unroll
            for (int q = 0; q < (RK * 27 + LBLOCK - 1) / LBLOCK; q++) {
                const int t = (int)threadIdx.x + q * LBLOCK;
                if (t < nv * 27) {
                    const int v = t / 27, e = t - v * 27;
                    int x, y, z;
                    kf_xyz(g, (long long)list_idx[ref_slot(pre, segcap, c0 + v)], x, y, z);
                    disp_at(u_in, g, x + e % 3 - 1, y + (e / 3) % 3 - 1, z + e / 9 - 1, s_nb[v][e]);
                }
            }
        }
        __syncthreads();
        const int total = nv * c.nsamp;
        for (int t = threadIdx.x; t < total; t += LBLOCK) {
            const int v = t / c.nsamp;
            const int si = t - v * c.nsamp;
            const int smp = si < c.nphi ? si : NPHI + (si - c.nphi);
            ref_sample<S, NB>(L, u_in, c.h, smp, s_xyz[v][0], s_xyz[v][1], s_xyz[v][2], s_f[v], &s_k[v][0][0],
                              s_nb[NB ? v : 0]);
        }
        __syncthreads();
        if ((int)threadIdx.x < nv) {
            const int t = threadIdx.x;
            float gr[3];
            ref_combine(c, s_psi[t], s_dst[t], s_f[t], s_k[t], gr);
            const int64_t j = s_i[t];
            if (MODE == KF_MODE_GRAD) {
                u_out[j] = gr[0];
                u_out[g.n + j] = gr[1];
                u_out[2 * g.n + j] = gr[2];
            } else {
                float d0 = -c.alpha * gr[0], d1 = -c.alpha * gr[1], d2 = -c.alpha * gr[2];
                bool fail = false;
                if (c.trust) {
                    // KillingFusion.cpp:288-314: shrink the step until the data residual at the
                    // moved sample, |src - dest|, is no larger than at the current one.  The
                    // comparison decides which voxels take a step and where a row stops, so it
                    // runs in the reference's FP64 (positions, reads and residuals): in FP32 the
                    // residual change of the smallest steps falls below one ulp and would read as
                    // "no larger" where FP64 keeps shrinking.
                    const int x = s_xyz[t][0], y = s_xyz[t][1], z = s_xyz[t][2];
                    const double ux = s_u[t][0], uy = s_u[t][1], uz = s_u[t][2];
                    const double dest = (double)s_dst[t];
                    const double at_u[3] = {((double)x + ux) + 0.5, ((double)y + uy) + 0.5, ((double)z + uz) + 0.5};
                    const double prev = L.distance64(at_u, c.oob64);
                    const double g0 = gr[0], g1 = gr[1], g2 = gr[2];
                    double a = c.alpha64;
                    double e0 = -a * g0, e1 = -a * g1, e2 = -a * g2;
                    for (;;) {
                        const double loc[3] = {(((double)x + ux) + e0) + 0.5, (((double)y + uy) + e1) + 0.5,
                                               (((double)z + uz) + e2) + 0.5};
                        const double src = L.distance64(loc, c.oob64);
                        if (!(fabs(src - dest) - fabs(prev - dest) > 0.0))
                            break;
                        a /= 1.5;
                        e0 = -a * g0;
                        e1 = -a * g1;
                        e2 = -a * g2;
                        if (!(a > 1e-7))
                            break;
                    }
                    d0 = (float)e0;
                    d1 = (float)e1;
                    d2 = (float)e2;
                    fail = a < 1e-7;
                    if (fail) {
                        // the reference's `break`: this voxel and the rest of its x-row take no step
                        atomicMin(&row_fail[(size_t)parity * (size_t)g.ny * g.nz + (size_t)z * g.ny + y],
                                  (unsigned int)x);
                        d0 = d1 = d2 = 0.0f;
                    }
                }
                float n0 = s_u[t][0] + d0, n1 = s_u[t][1] + d1, n2 = s_u[t][2] + d2;
                u_out[j] = n0;
                u_out[g.n + j] = n1;
                u_out[2 * g.n + j] = n2;
                double dn = sqrt((double)d0 * d0 + (double)d1 * d1 + (double)d2 * d2);
                if (c.trust) {
                    // the statistics and the divergence test wait for the row's first failure
                    // (k_ref_trust_rows); the entry's warped value is used up
                    list_psi[s_q[t]] = fail ? -1.0f : (float)dn;
                } else if (!(isfinite(n0) && isfinite(n1) && isfinite(n2))) {
                    atomicMin((long long *)&st->div_slot[parity], (long long)j);
                } else if (dn > mx) {
                    mx = dn;
                }
            }
        }
        __syncthreads();
    }
    if (MODE == KF_MODE_STEP && !c.trust)
        kf_block_publish<LBLOCK>(partials, blockIdx.x, 0.0, mx, blockIdx.x == 0 && threadIdx.x == 0 ? count : 0ull);
}

// Pass 3 of a trust-region step: the reference's `break` (KillingFusion.cpp:313) leaves the rest
// of an x-row unvisited once a voxel's step fails, so every listed voxel past its row's first
// failure (row_fail, set by pass 2) takes no step after all, u_out = u_in + the delta field's
// +0.  The statistics are the reference loop's: the voxels it visited in the band (up to and
// including each row's failure) and the largest step it kept; the divergence test covers the
// kept steps.  The row failures of the other launch parity are cleared for the next launch.
__global__ __launch_bounds__(LBLOCK) void k_ref_trust_rows(GridDims g, const float *__restrict__ u_in,
                                                           float *__restrict__ u_out, DeformState *st,
                                                           BlockPartial *partials, const unsigned long long *aux,
                                                           int parity, unsigned long long segcap,
                                                           const unsigned long long *__restrict__ list_idx,
                                                           const float *__restrict__ list_dn,
                                                           unsigned int *__restrict__ row_fail)
{
    if (st->done)
        return;
    const size_t rows = (size_t)g.ny * g.nz;
    unsigned int *fail_now = row_fail + (size_t)parity * rows;
    unsigned int *fail_next = row_fail + (size_t)(parity ^ 1) * rows;
    for (size_t r = (size_t)blockIdx.x * LBLOCK + threadIdx.x; r < rows; r += (size_t)gridDim.x * LBLOCK)
        fail_next[r] = 0xffffffffu;
    unsigned long long pre[NSEG + 1];
    pre[0] = 0ull;
#pragma unroll
    for (int k = 0; k < NSEG; k++)
        pre[k + 1] = pre[k] + aux[parity * NSEG + k];
    const unsigned long long count = pre[NSEG];
    double mx = 0.0;
    unsigned long long act = 0;
    for (unsigned long long p = (unsigned long long)blockIdx.x * LBLOCK + threadIdx.x; p < count;
         p += (unsigned long long)gridDim.x * LBLOCK) {
        const unsigned long long q = ref_slot(pre, segcap, p);
        const long long j = (long long)list_idx[q];
        int x, y, z;
        kf_xyz(g, j, x, y, z);
        const unsigned int f = fail_now[(size_t)z * g.ny + y];
        if ((unsigned int)x > f) {
#pragma unroll
            for (int a = 0; a < 3; a++)
                u_out[a * g.n + j] = u_in[a * g.n + j] + 0.0f;
            continue;
        }
        act += 1;
        if ((unsigned int)x == f)
            continue;  // the failing voxel: no step (pass 2 wrote u + 0)
        const float n0 = u_out[j], n1 = u_out[g.n + j], n2 = u_out[2 * g.n + j];
        if (!(isfinite(n0) && isfinite(n1) && isfinite(n2)))
            atomicMin((long long *)&st->div_slot[parity], (long long)j);
        else if ((double)list_dn[q] > mx)
            mx = (double)list_dn[q];
    }
    kf_block_publish<LBLOCK>(partials, blockIdx.x, 0.0, mx, act);
}

__global__ void k_ref_rows_clear(unsigned int *row_fail, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        row_fail[i] = 0xffffffffu;
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

void kf_ref_release(kf_context *ctx)
{
    kf_ref_ws &w = ctx->ref;
    if (w.idx)
        (void)hipFree(w.idx);
    if (w.psi)
        (void)hipFree(w.psi);
    if (w.aux)
        (void)hipFree(w.aux);
    if (w.row_fail)
        (void)hipFree(w.row_fail);
    w = kf_ref_ws();
}

kf_status kf_launch_ref_step(const kf_volume *live, const kf_volume *can, const float *u_in, float *u_out,
                             const kf_params *p, int mode, int iter, const unsigned int *vlist, unsigned int vslots)
{
    kf_context *ctx = live->ctx;
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    const int parity = mode == KF_MODE_GRAD ? 0 : (iter & 1);
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
    c.nphi = c.use_ls ? NPHI : (c.use_data ? 6 : 0);
    c.nsamp = c.nphi + (c.use_k ? NKIL : 0);
    // KillingFusion.cpp:288: the trust region only with the data energy alone
    c.trust = mode != KF_MODE_GRAD && p->use_trust && c.use_data && !c.use_ls && !c.use_k;
    c.alpha64 = p->alpha;
    c.oob64 = live->oob();
    const int64_t n = live->n;
    const bool vec = n % 4 == 0;
    const bool nbhd = c.h > 0.0f && c.h < 0.25f;
    const int64_t per_block = 4 * SBLOCK;
    const unsigned int nb_a = (unsigned int)((n + per_block - 1) / per_block);
    // Segment x takes the band voxels of the pass-1 blocks b with b % 8 == x: at most their
    // voxels, so every segment has room for ceil(nb_a / 8) blocks.
    const unsigned long long segcap = (unsigned long long)((nb_a + NSEG - 1) / NSEG) * (unsigned long long)per_block;
    const size_t need = (size_t)segcap * NSEG;
    kf_ref_ws &w = ctx->ref;
    if (w.cap < need) {
        KF_HIP(hipStreamSynchronize(ctx->stream));  // earlier launches may still read the old list
        kf_ref_release(ctx);
        if (hipMalloc((void **)&w.idx, sizeof(unsigned long long) * need) != hipSuccess ||
            hipMalloc((void **)&w.psi, sizeof(float) * need) != hipSuccess ||
            hipMalloc((void **)&w.aux, AUX_WORDS * sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            kf_ref_release(ctx);
            kf_set_error("reference-step list allocation failed");
            return KF_E_OOM;
        }
        w.cap = need;
    }
    // The active-set pass 1 runs blocks of SBLOCK list slots (vslots of them, the list's length
    // read on the device); its segments then hold at most ceil(blocks / 8) blocks' voxels each,
    // within the capacity sized for the whole grid.
    const unsigned int nb_v = vlist ? (vslots + SBLOCK - 1) / SBLOCK : 0u;
    const unsigned long long segcap_v = (unsigned long long)((nb_v + NSEG - 1) / NSEG) * SBLOCK;
    KF_CHECK_ARG(!vlist || (mode != KF_MODE_GRAD && segcap_v * NSEG <= w.cap), "active list larger than the workspace");
    // the first failing x of every (y, z) row, one set per launch parity
    const size_t rows = (size_t)live->dims[1] * (size_t)live->dims[2];
    if (c.trust && w.rows_cap < 2 * rows) {
        KF_HIP(hipStreamSynchronize(ctx->stream));
        if (w.row_fail)
            (void)hipFree(w.row_fail);
        w.row_fail = nullptr;
        w.rows_cap = 0;
        if (hipMalloc((void **)&w.row_fail, sizeof(unsigned int) * 2 * rows) != hipSuccess) {
            (void)hipGetLastError();
            kf_set_error("trust-region row table allocation failed");
            return KF_E_OOM;
        }
        w.rows_cap = 2 * rows;
    }
    kf_status s = KF_OK;
    const int64_t chunks = (n + RK - 1) / RK;
    const unsigned int nb_b = (unsigned int)(chunks < LIST_BLOCKS ? chunks : LIST_BLOCKS);
    if (mode != KF_MODE_GRAD && (s = kf_ensure_partials(ctx, nb_b)) != KF_OK)
        return s;
    BlockPartial *parts = kf_partials(ctx, parity);
    KF_DISPATCH_DTYPE(live->dtype, S, {
        PhiView<S> L = make_view<S>(live), Cn = make_view<S>(can);
        if (mode == KF_MODE_GRAD || iter == 0) {
            // A deform call (or a gradient evaluation) starts with empty lists and checks the live
            // SDF once: it does not change while u does.
            hipLaunchKernelGGL(k_ref_clear, dim3(1), dim3(64), 0, ctx->stream, w.aux);
            if (c.trust)
                hipLaunchKernelGGL(k_ref_rows_clear, dim3(64), dim3(256), 0, ctx->stream, w.row_fail, 2 * rows);
            const unsigned int fb = nb_a < 4096u ? nb_a : 4096u;
            if (vec)
                hipLaunchKernelGGL((k_ref_finite<S, 4>), dim3(fb), dim3(BLOCK), 0, ctx->stream, L, w.aux + AUX_NONFINITE);
            else
                hipLaunchKernelGGL((k_ref_finite<S, 1>), dim3(fb), dim3(BLOCK), 0, ctx->stream, L, w.aux + AUX_NONFINITE);
        }
        if (mode == KF_MODE_GRAD) {
            if (vec)
                hipLaunchKernelGGL((k_ref_scan<S, KF_MODE_GRAD, false>), dim3(nb_a), dim3(SBLOCK), 0, ctx->stream, L,
                                   u_in, u_out, c.hi, c.lo, c.alpha, ctx->state, w.aux, parity, segcap, w.idx, w.psi);
            else
                hipLaunchKernelGGL((k_ref_scan<S, KF_MODE_GRAD, true>), dim3(nb_a), dim3(SBLOCK), 0, ctx->stream, L,
                                   u_in, u_out, c.hi, c.lo, c.alpha, ctx->state, w.aux, parity, segcap, w.idx, w.psi);
            if (nbhd)
                hipLaunchKernelGGL((k_ref_list<S, KF_MODE_GRAD, true>), dim3(nb_b), dim3(LBLOCK), 0, ctx->stream, L,
                                   Cn, u_in, u_out, c, ctx->state, parts, w.aux, parity, segcap, w.idx, w.psi, w.row_fail);
            else
                hipLaunchKernelGGL((k_ref_list<S, KF_MODE_GRAD, false>), dim3(nb_b), dim3(LBLOCK), 0, ctx->stream, L,
                                   Cn, u_in, u_out, c, ctx->state, parts, w.aux, parity, segcap, w.idx, w.psi, w.row_fail);
        } else {
            const unsigned long long sc = vlist ? segcap_v : segcap;
            if (vlist) {
                if (nb_v > 0)
                    hipLaunchKernelGGL((k_ref_active_scan<S, KF_MODE_STEP>), dim3(nb_v), dim3(SBLOCK), 0, ctx->stream,
                                       L, u_in, u_out, c.hi, c.lo, c.alpha, ctx->state, w.aux, parity, sc, w.idx,
                                       w.psi, vlist);
            } else if (vec) {
                hipLaunchKernelGGL((k_ref_scan<S, KF_MODE_STEP, false>), dim3(nb_a), dim3(SBLOCK), 0, ctx->stream, L,
                                   u_in, u_out, c.hi, c.lo, c.alpha, ctx->state, w.aux, parity, sc, w.idx, w.psi);
            } else {
                hipLaunchKernelGGL((k_ref_scan<S, KF_MODE_STEP, true>), dim3(nb_a), dim3(SBLOCK), 0, ctx->stream, L,
                                   u_in, u_out, c.hi, c.lo, c.alpha, ctx->state, w.aux, parity, sc, w.idx, w.psi);
            }
            if (nbhd)
                hipLaunchKernelGGL((k_ref_list<S, KF_MODE_STEP, true>), dim3(nb_b), dim3(LBLOCK), 0, ctx->stream, L,
                                   Cn, u_in, u_out, c, ctx->state, parts, w.aux, parity, sc, w.idx, w.psi, w.row_fail);
            else
                hipLaunchKernelGGL((k_ref_list<S, KF_MODE_STEP, false>), dim3(nb_b), dim3(LBLOCK), 0, ctx->stream, L,
                                   Cn, u_in, u_out, c, ctx->state, parts, w.aux, parity, sc, w.idx, w.psi, w.row_fail);
        }
    });
    KF_LAUNCH_CHECK();
    if (c.trust) {
        const unsigned long long sc = vlist ? segcap_v : segcap;
        hipLaunchKernelGGL(k_ref_trust_rows, dim3(nb_b), dim3(LBLOCK), 0, ctx->stream, grid_of(live), u_in, u_out,
                           ctx->state, parts, (const unsigned long long *)w.aux, parity, sc,
                           (const unsigned long long *)w.idx, (const float *)w.psi, w.row_fail);
        KF_LAUNCH_CHECK();
    }
    return mode == KF_MODE_GRAD ? KF_OK : kf_launch_finalize(ctx, nb_b, parity);
}
