// k_rigid.hip -- rigid SDF-to-SDF pre-registration (no reference counterpart: the reference
// leaves it as a future task, KillingFusion.cpp:76,133), and the displacement field of a
// rigid pose, which starts the non-rigid optimiser from the rigid alignment.
//
// Pose T maps canonical world points to live world points (the direction of a displacement
// field: canonical voxel j samples the live SDF at j + u_j).  Gauss-Newton on
//   E(T) = 1/2 sum_j (phi_live(T x_j) - phi_can(x_j))^2
// over canonical voxels j with weight > 0 and -clip <= phi_can <= Max - eps whose sample
// T x_j falls inside the live grid's interpolation domain.  phi_live is the trilinear
// interpolant over the live nodes (node k at world min + voxel (k + 1/2), the centre the
// integration assigns to voxel k).  With the left perturbation T <- (I + [w]x) T + tau,
// about the grid centre c (x' = T x - c):  J = [x' x g, g],  g = grad phi_live at T x.
// Each block accumulates J^T J (21 values), J^T r (6), sum r^2 and the count in FP64;
// one block folds the partials in block order, so the normal equations are reproducible.
#include "kf_reduce.hpp"

#include <cmath>
#include <cstring>

namespace {

constexpr int BLOCK = 256;
constexpr int NSUM = 29;  // 21 (upper triangle of J^T J) + 6 (J^T r) + sum r^2 + count

struct Pose {
    double R[9];  // row-major
    double t[3];
};

template <class S>
struct RigidView {
    const typename S::T *phi;
    const int32_t *w;
    GridDims g;
    float scale;
    double min[3], voxel;
};

template <class S>
__device__ inline double node(const RigidView<S> &V, int x, int y, int z)
{
    return (double)S::load(V.phi, V.g.idx(x, y, z), V.scale);
}

// Trilinear value and index-space gradient at index coordinates (qx, qy, qz); false outside
// the interpolation domain [0, n - 1]^3.
template <class S>
__device__ inline bool sample(const RigidView<S> &V, double qx, double qy, double qz, double &val, double grad[3])
{
    if (V.g.nx < 2 || V.g.ny < 2 || V.g.nz < 2)
        return false;
    if (!(qx >= 0.0 && qy >= 0.0 && qz >= 0.0 && qx <= V.g.nx - 1 && qy <= V.g.ny - 1 && qz <= V.g.nz - 1))
        return false;
    const int bx = min((int)qx, V.g.nx - 2), by = min((int)qy, V.g.ny - 2), bz = min((int)qz, V.g.nz - 2);
    const double fx = qx - bx, fy = qy - by, fz = qz - bz;
    double c[8];
    for (int k = 0; k < 8; k++)
        c[k] = node(V, bx + (k & 1), by + ((k >> 1) & 1), bz + (k >> 2));
    const double x00 = c[0] + fx * (c[1] - c[0]), x10 = c[2] + fx * (c[3] - c[2]);
    const double x01 = c[4] + fx * (c[5] - c[4]), x11 = c[6] + fx * (c[7] - c[6]);
    const double y0 = x00 + fy * (x10 - x00), y1 = x01 + fy * (x11 - x01);
    val = y0 + fz * (y1 - y0);
    const double dx00 = c[1] - c[0], dx10 = c[3] - c[2], dx01 = c[5] - c[4], dx11 = c[7] - c[6];
    const double dxy0 = dx00 + fy * (dx10 - dx00), dxy1 = dx01 + fy * (dx11 - dx01);
    grad[0] = dxy0 + fz * (dxy1 - dxy0);
    const double dy0 = (c[2] + fx * (c[3] - c[2])) - (c[0] + fx * (c[1] - c[0]));
    const double dy1 = (c[6] + fx * (c[7] - c[6])) - (c[4] + fx * (c[5] - c[4]));
    grad[1] = dy0 + fz * (dy1 - dy0);
    grad[2] = y1 - y0;
    return true;
}

template <class S>
__global__ __launch_bounds__(BLOCK) void k_rigid_normal_eq(RigidView<S> live, RigidView<S> can, Pose T, double cx,
                                                           double cy, double cz, double hi, double lo,
                                                           double *partials)
{
    double acc[NSUM];
#pragma unroll
    for (int k = 0; k < NSUM; k++)
        acc[k] = 0.0;
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j < can.g.n && can.w[j] > 0) {
        const double pc = node(can, (int)(j % can.g.nx), (int)((j / can.g.nx) % can.g.ny),
                               (int)(j / ((int64_t)can.g.nx * can.g.ny)));
        if (pc >= lo && pc <= hi) {
            const int x = (int)(j % can.g.nx), y = (int)((j / can.g.nx) % can.g.ny);
            const int z = (int)(j / ((int64_t)can.g.nx * can.g.ny));
            const double X[3] = {can.min[0] + can.voxel * (x + 0.5), can.min[1] + can.voxel * (y + 0.5),
                                 can.min[2] + can.voxel * (z + 0.5)};
            double Y[3];
            for (int r = 0; r < 3; r++)
                Y[r] = T.R[3 * r] * X[0] + T.R[3 * r + 1] * X[1] + T.R[3 * r + 2] * X[2] + T.t[r];
            const double q[3] = {(Y[0] - live.min[0]) / live.voxel - 0.5, (Y[1] - live.min[1]) / live.voxel - 0.5,
                                 (Y[2] - live.min[2]) / live.voxel - 0.5};
            double val, gi[3];
            if (sample(live, q[0], q[1], q[2], val, gi)) {
                const double g[3] = {gi[0] / live.voxel, gi[1] / live.voxel, gi[2] / live.voxel};
                const double p[3] = {Y[0] - cx, Y[1] - cy, Y[2] - cz};
                const double J[6] = {p[1] * g[2] - p[2] * g[1], p[2] * g[0] - p[0] * g[2], p[0] * g[1] - p[1] * g[0],
                                     g[0], g[1], g[2]};
                const double res = val - pc;
                int k = 0;
                for (int a = 0; a < 6; a++)
                    for (int b = a; b < 6; b++)
                        acc[k++] = J[a] * J[b];
                for (int a = 0; a < 6; a++)
                    acc[21 + a] = J[a] * res;
                acc[27] = res * res;
                acc[28] = 1.0;
            }
        }
    }
    // Block sums in a fixed order (wavefront butterflies, then waves in order).
    __shared__ double s[BLOCK / 64][NSUM];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NSUM; k++) {
        const double v = kf_wave_sum(acc[k]);
        if (lane == 0)
            s[wid][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NSUM) {
        double v = 0.0;
        for (int w = 0; w < BLOCK / 64; w++)
            v += s[w][threadIdx.x];
        partials[(size_t)blockIdx.x * NSUM + threadIdx.x] = v;
    }
}

// Folds nb block partials into out[NSUM] (one block; thread k owns sum k, blocks in order).
__global__ __launch_bounds__(64) void k_rigid_fold(const double *partials, unsigned int nb, double *out)
{
    if (threadIdx.x < NSUM) {
        double v = 0.0;
        for (unsigned int b = 0; b < nb; b++)
            v += partials[(size_t)b * NSUM + threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// u_j = T(j) - j in live index units for every canonical voxel j (a rigid displacement field).
__global__ __launch_bounds__(BLOCK) void k_field_from_pose(float *u, GridDims g, double cmin0, double cmin1,
                                                           double cmin2, double cvox, double lmin0, double lmin1,
                                                           double lmin2, double lvox, Pose T)
{
    const int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= g.n)
        return;
    const int x = (int)(j % g.nx), y = (int)((j / g.nx) % g.ny), z = (int)(j / ((int64_t)g.nx * g.ny));
    const double X[3] = {cmin0 + cvox * (x + 0.5), cmin1 + cvox * (y + 0.5), cmin2 + cvox * (z + 0.5)};
    const double lm[3] = {lmin0, lmin1, lmin2};
    const int idx[3] = {x, y, z};
    for (int r = 0; r < 3; r++) {
        const double Y = T.R[3 * r] * X[0] + T.R[3 * r + 1] * X[1] + T.R[3 * r + 2] * X[2] + T.t[r];
        u[(int64_t)r * g.n + j] = (float)((Y - lm[r]) / lvox - 0.5 - idx[r]);
    }
}

template <class S>
RigidView<S> view_of(const kf_volume *v)
{
    RigidView<S> V;
    V.phi = (const typename S::T *)v->phi;
    V.w = v->w;
    V.g = grid_of(v);
    V.scale = v->scale();
    for (int a = 0; a < 3; a++)
        V.min[a] = v->min_corner[a];
    V.voxel = v->voxel;
    return V;
}

Pose pose_from16(const double m[16])
{
    Pose P;
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++)
            P.R[3 * r + c] = m[4 * r + c];
        P.t[r] = m[4 * r + 3];
    }
    return P;
}

// Solves the 6 x 6 symmetric positive system A x = b (upper triangle packed row-wise) by
// Cholesky; false if A is not positive definite.
bool solve6(const double up[21], const double b[6], double x[6])
{
    double A[6][6];
    int k = 0;
    for (int i = 0; i < 6; i++)
        for (int j = i; j < 6; j++)
            A[i][j] = A[j][i] = up[k++];
    double L[6][6] = {};
    for (int i = 0; i < 6; i++) {
        for (int j = 0; j <= i; j++) {
            double s = A[i][j];
            for (int m = 0; m < j; m++)
                s -= L[i][m] * L[j][m];
            if (i == j) {
                if (!(s > 0.0))
                    return false;
                L[i][i] = std::sqrt(s);
            } else {
                L[i][j] = s / L[j][j];
            }
        }
    }
    double yv[6];
    for (int i = 0; i < 6; i++) {
        double s = b[i];
        for (int m = 0; m < i; m++)
            s -= L[i][m] * yv[m];
        yv[i] = s / L[i][i];
    }
    for (int i = 5; i >= 0; i--) {
        double s = yv[i];
        for (int m = i + 1; m < 6; m++)
            s -= L[m][i] * x[m];
        x[i] = s / L[i][i];
    }
    return true;
}

// Rodrigues rotation of vector w (angle |w|).
void rotation(const double w[3], double R[9])
{
    const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double K2[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int m = 0; m < 3; m++)
                s += K[3 * r + m] * K[3 * m + c];
            K2[3 * r + c] = s;
        }
    const double a = th > 1e-12 ? std::sin(th) / th : 1.0 - th * th / 6.0;
    const double b = th > 1e-12 ? (1.0 - std::cos(th)) / (th * th) : 0.5 - th * th / 24.0;
    for (int i = 0; i < 9; i++)
        R[i] = (i % 4 == 0 ? 1.0 : 0.0) + a * K[i] + b * K2[i];
}

}  // namespace

extern "C" {

kf_status kf_rigid_align(const kf_volume *live, const kf_volume *can, int32_t max_iters, double pose[16],
                         kf_rigid_stats *stats)
{
    KF_CHECK_ARG(live && can && pose && max_iters >= 0, "null argument");
    KF_CHECK_ARG(live->ctx == can->ctx && live->dtype == can->dtype, "volumes need one context and dtype");
    kf_context *ctx = can->ctx;
    KfDeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    const unsigned int nb = (unsigned int)((can->n + BLOCK - 1) / BLOCK);
    double *dpart = nullptr, *dsum = nullptr;
    KF_HIP(hipMallocAsync((void **)&dpart, sizeof(double) * ((size_t)nb + 1) * NSUM, st));
    dsum = dpart + (size_t)nb * NSUM;
    const double cx = can->min_corner[0] + 0.5 * can->voxel * can->dims[0];
    const double cy = can->min_corner[1] + 0.5 * can->voxel * can->dims[1];
    const double cz = can->min_corner[2] + 0.5 * can->voxel * can->dims[2];
    const double hi = can->max_surface - can->epsilon, lo = -can->clip;
    double h[NSUM];
    kf_rigid_stats out{};
    kf_status s = KF_OK;
    double T[16];
    std::memcpy(T, pose, sizeof(T));
    int it = 0;
    for (; it <= max_iters && s == KF_OK; it++) {
        const Pose P = pose_from16(T);
        KF_DISPATCH_DTYPE(can->dtype, S, {
            hipLaunchKernelGGL((k_rigid_normal_eq<S>), dim3(nb), dim3(BLOCK), 0, st, view_of<S>(live), view_of<S>(can),
                               P, cx, cy, cz, hi, lo, dpart);
        });
        hipLaunchKernelGGL(k_rigid_fold, dim3(1), dim3(64), 0, st, (const double *)dpart, nb, dsum);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(h, dsum, sizeof(h), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            kf_set_error(std::string("rigid alignment: ") + hipGetErrorString(e));
            s = KF_E_HIP;
            break;
        }
        out.energy = 0.5 * h[27];
        out.voxels = (uint64_t)h[28];
        if (it == 0)
            out.initial_energy = out.energy;
        if (it == max_iters || h[28] < 6)
            break;
        double b[6], x[6];
        for (int a = 0; a < 6; a++)
            b[a] = -h[21 + a];
        if (!solve6(h, b, x))
            break;
        // T <- exp([w]x) about the centre, then translation: Y' = R (Y - c) + c + tau
        double R[9];
        rotation(x, R);
        double N[16];
        const double c3[3] = {cx, cy, cz};
        for (int r = 0; r < 3; r++) {
            for (int col = 0; col < 4; col++) {
                double acc = 0.0;
                for (int m = 0; m < 3; m++)
                    acc += R[3 * r + m] * T[4 * m + col];
                N[4 * r + col] = acc;
            }
            double rc = 0.0;
            for (int m = 0; m < 3; m++)
                rc += R[3 * r + m] * c3[m];
            N[4 * r + 3] += c3[r] - rc + x[3 + r];
        }
        N[12] = N[13] = N[14] = 0.0;
        N[15] = 1.0;
        std::memcpy(T, N, sizeof(T));
        out.iters = it + 1;
        out.last_step = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) +
                        std::sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]) / can->voxel;
        if (out.last_step < 1e-9)
            break;
    }
    (void)hipFreeAsync(dpart, st);
    if (s != KF_OK)
        return s;
    std::memcpy(pose, T, sizeof(T));
    if (stats)
        *stats = out;
    return KF_OK;
}

kf_status kf_field_from_rigid(kf_field *u, const kf_volume *can, const kf_volume *live, const double pose[16])
{
    KF_CHECK_ARG(u && can && live && pose, "null argument");
    KF_CHECK_ARG(same_dims(u->dims, can->dims), "field and canonical volume dims differ");
    KfDeviceGuard dg(u->ctx->device);
    const Pose P = pose_from16(pose);
    hipLaunchKernelGGL(k_field_from_pose, dim3((unsigned int)((u->n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                       u->ctx->stream, u->u(), grid_of(u), can->min_corner[0], can->min_corner[1], can->min_corner[2],
                       can->voxel, live->min_corner[0], live->min_corner[1], live->min_corner[2], live->voxel, P);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

}  // extern "C"
