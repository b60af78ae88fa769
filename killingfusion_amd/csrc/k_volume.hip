// k_volume.hip -- per-frame volume kernels: fill, storage conversion, projective
// integration, warp, fuse, displacement-aware fuse, field fill/add.
//
// These kernels make discrete decisions (pixel pick, interpolated weights truncated to
// integers) exactly where the reference does, so they compute in FP64 with the reference's
// operation order, and this file is built with -ffp-contract=off: every product and sum
// rounds on its own, as in the reference's x86-64 build.  They stream each voxel once.
#include "kf_internal.hpp"
#include "kf_project.hpp"

namespace {

constexpr int BLOCK = 256;

inline unsigned int blocks_for(int64_t n) { return (unsigned int)((n + BLOCK - 1) / BLOCK); }

// SDF.cpp:133-134: phi = Max + eps, w = 0.  Four voxels per thread, phi and w each as one
// streaming vector store (the volume is far larger than the L2).
template <class S>
__global__ void k_fill(typename S::T *phi, int32_t *w, int64_t n, typename S::T init)
{
    const int64_t i0 = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * 4;
    if (i0 + 4 <= n) {
        typedef typename S::T vec_t __attribute__((ext_vector_type(4)));
        typedef int32_t ivec_t __attribute__((ext_vector_type(4)));
        const vec_t pv = {init, init, init, init};
        const ivec_t wv = {0, 0, 0, 0};
        __builtin_nontemporal_store(pv, reinterpret_cast<vec_t *>(phi + i0));
        __builtin_nontemporal_store(wv, reinterpret_cast<ivec_t *>(w + i0));
    } else {
        for (int64_t i = i0; i < n; i++) {
            phi[i] = init;
            w[i] = 0;
        }
    }
}

template <class S>
__global__ void k_encode(typename S::T *phi, const float *src, int64_t n, float scale)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n)
        phi[i] = S::enc(src[i], scale);
}

template <class S>
__global__ void k_decode(float *dst, const typename S::T *phi, int64_t n, float scale)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n)
        dst[i] = S::load(phi, i, scale);
}

// SDF::integrateDepthFrame, SDF.cpp:149-196.  FRESH: the same pass also resets the volume
// (SDF.cpp:133-134: phi = Max + eps, w = 0) before integrating, as a reset followed by an
// integration would -- same values, one write per voxel and no read.
//
// A thread takes IPT consecutive voxels.  Each voxel's pixel is decided by kf_pixel_fast (no
// FP64 division), or by kf_pixel_rowwise for cameras whose depth and image row stay fixed along
// a grid row (ROWWISE: the reference's identity pose), and, where that route cannot vouch for
// the reference's rounding, by kf_pixel_exact (kf_project.hpp); the camera depth c2 is the
// reference's FP64 value either way.  The running average (old w + dist) / (w + 1) is dist
// itself when w = 0 and old is finite (old 0 = +-0, +-0 + dist = dist for dist != 0 and +0 for
// dist = +0, dist / 1 = dist), so that case skips the multiply, add and FP64 division.  The
// stores of a FRESH pass are one streaming vector store per thread for phi and for w.
constexpr int IPT = 4;

template <class S>
struct IntegrateArgs {
    typename S::T *phi;
    int32_t *w;
    GridDims g;
    double mn[3];
    double voxel, clip;
    float scale;
    double rscale;          // kf_quot_rscale(scale)
    typename S::T init;     // the stored Max + eps of a voxel no frame has reached
    const float *__restrict__ depth;
    KfCamera C;
    double min_depth, max_depth;
};

// Per-row constants: the exact camera-depth products of row (y, z) and the fast path's sums.
struct IntegrateRow {
    double p2[3];
    double bx, by;
};

template <class S>
__device__ inline IntegrateRow integrate_row(const IntegrateArgs<S> &A, int y, int z)
{
    const double Z = A.voxel * (z + 0.5) + A.mn[2];
    const double Y = A.voxel * (y + 0.5) + A.mn[1];
    IntegrateRow R;
    R.p2[0] = A.C.M[9] * Y;
    R.p2[1] = A.C.M[10] * Z;
    R.p2[2] = A.C.M[11] * 1.0;
    kf_row_fast(A.C, Y, Z, R.bx, R.by);
    return R;
}

template <class S>
__device__ inline KfRowwise integrate_rowwise(const IntegrateArgs<S> &A, int y, int z)
{
    const double Z = A.voxel * (z + 0.5) + A.mn[2];
    const double Y = A.voxel * (y + 0.5) + A.mn[1];
    const double c2 = kf_rowwise_depth(A.C, Y, Z);
    return kf_rowwise_row(A.C, Y, Z, c2, __builtin_amdgcn_rcpf((float)c2), A.voxel, A.mn[0]);
}

// A voxel's camera depth and the index of the pixel it reads, or -1 outside the image.
struct Projected {
    double zc;
    int pix;
};

// The running average of SDF.cpp:189-193 for a voxel that read depth d (pix >= 0); false when
// the frame leaves the voxel alone (depth out of range, or too far behind the surface).
template <class S>
__device__ inline bool average_voxel(const IntegrateArgs<S> &A, const Projected &pr, float depth, int32_t wi,
                                     double old, double &nv)
{
    if (pr.pix < 0)
        return false;
    const double d = (double)depth;
    if (d < A.min_depth || d > A.max_depth)
        return false;
    const double dist = d - pr.zc;
    if (dist < -A.clip)
        return false;
    nv = wi == 0 && isfinite(old) ? dist : (old * (double)wi + dist) / (double)(wi + 1);
    return true;
}

// ALIGNED: rows hold a multiple of IPT voxels, so a thread's group lies in one row and inside
// the grid (192^3 BF16 update pass: 12.9-13.1 against 14.3-14.4 us at the minimum).
template <class S, bool FRESH, bool ROWWISE, bool ALIGNED>
__global__ __launch_bounds__(BLOCK) void k_integrate(IntegrateArgs<S> A)
{
    const GridDims &g = A.g;
    const int64_t i0 = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * IPT;
    if (i0 >= g.n)
        return;
    const bool whole = ALIGNED || i0 + IPT <= g.n;
    int x, y, z;
    kf_xyz(g, i0, x, y, z);
    // project the group's voxels, then read their pixels together
    Projected pr[IPT];
    int vx[IPT], vy[IPT], vz[IPT];
    unsigned int exact = 0;  // voxels whose pixel kf_pixel_exact decides
    if (ROWWISE) {
        // depth and image row fixed along the grid row (kf_camera_rowwise)
        KfRowwise R = integrate_rowwise(A, y, z);
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            if (!ALIGNED && x == g.nx) {  // the group runs into the next row
                x = 0;
                if (++y == g.ny) {
                    y = 0;
                    ++z;
                }
                R = integrate_rowwise(A, y, z);
            }
            bool ok;
            const int pix = kf_pixel_rowwise(A.C, R, x, ok);
            const bool live = ALIGNED || i0 + k < g.n;
            pr[k] = Projected{R.c2, live ? pix : -1};
            exact |= (live && !ok ? 1u : 0u) << k;
            vx[k] = x;
            vy[k] = y;
            vz[k] = z;
            ++x;
        }
    } else {
        IntegrateRow R = integrate_row(A, y, z);
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            if (!ALIGNED && x == g.nx) {
                x = 0;
                if (++y == g.ny) {
                    y = 0;
                    ++z;
                }
                R = integrate_row(A, y, z);
            }
            const double X = A.voxel * (x + 0.5) + A.mn[0];
            const double c2 = ((A.C.M[8] * X + R.p2[0]) + R.p2[1]) + R.p2[2];
            bool ok;
            const int pix = kf_pixel_fast(A.C, R.bx, R.by, X, c2, __builtin_amdgcn_rcpf((float)c2), ok);
            const bool live = ALIGNED || i0 + k < g.n;
            pr[k] = Projected{c2, live ? pix : -1};
            exact |= (live && !ok ? 1u : 0u) << k;
            vx[k] = x;
            vy[k] = y;
            vz[k] = z;
            ++x;
        }
    }
    if (exact) {
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            if (exact >> k & 1u) {
                double P[3][3];
                kf_row_exact(A.C, A.voxel * (vy[k] + 0.5) + A.mn[1], A.voxel * (vz[k] + 0.5) + A.mn[2], P);
                const double X = A.voxel * (vx[k] + 0.5) + A.mn[0];
                pr[k].zc = kf_cam(A.C, P[2], 2, X);
                pr[k].pix = kf_pixel_exact(A.C, P, X, pr[k].zc);
            }
        }
    }
    float dep[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++)
        dep[k] = A.depth[pr[k].pix >= 0 ? pr[k].pix : 0];
    if (FRESH && whole) {
        // all IPT voxels written: phi and w as one vector store each
        typedef typename S::T vec_t __attribute__((ext_vector_type(IPT)));
        vec_t pv;
        int32_t wv[IPT];
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            double nv;
            const bool take = average_voxel(A, pr[k], dep[k], 0, 0.0, nv);
            pv[k] = take ? S::enc_r((float)nv, A.scale, A.rscale) : A.init;
            wv[k] = take ? 1 : 0;
        }
        typedef int32_t ivec_t __attribute__((ext_vector_type(IPT)));
        __builtin_nontemporal_store(pv, reinterpret_cast<vec_t *>(A.phi + i0));
        __builtin_nontemporal_store(ivec_t{wv[0], wv[1], wv[2], wv[3]}, reinterpret_cast<ivec_t *>(A.w + i0));
        return;
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int64_t i = i0 + k;
        if (i >= g.n)
            break;
        const int32_t wi = FRESH ? 0 : A.w[i];
        const double old = FRESH ? (double)S::dec(A.init, A.scale) : (double)S::load(A.phi, i, A.scale);
        double nv;
        if (average_voxel(A, pr[k], dep[k], wi, old, nv)) {
            A.phi[i] = S::enc_r((float)nv, A.scale, A.rscale);
            A.w[i] = wi + 1;
        } else if (FRESH) {
            A.phi[i] = A.init;
            A.w[i] = 0;
        }
    }
}

template <class S>
struct Reader {
    const typename S::T *phi;
    const int32_t *w;
    GridDims g;
    float scale;
    double oob;
    __device__ inline double dist(int x, int y, int z) const
    {
        return g.in(x, y, z) ? (double)S::load(phi, g.idx(x, y, z), scale) : oob;
    }
    __device__ inline double weight(int x, int y, int z) const
    {
        return g.in(x, y, z) ? (double)w[g.idx(x, y, z)] : 0.0;
    }
    // SDF::getDistance / getWeight(Vector3d), SDF.cpp:503-524, 584-604, in FP64.
    __device__ inline void both(double lx, double ly, double lz, double &dd, double &ww) const
    {
        const double tx = lx - 0.5, ty = ly - 0.5, tz = lz - 0.5;
        const int bx = (int)tx, by = (int)ty, bz = (int)tz;
        const double fx = tx - (double)bx, fy = ty - (double)by, fz = tz - (double)bz;
        ww = kf_interp3(weight(bx, by, bz), weight(bx + 1, by, bz), weight(bx, by + 1, bz),
                        weight(bx + 1, by + 1, bz), weight(bx, by, bz + 1), weight(bx + 1, by, bz + 1),
                        weight(bx, by + 1, bz + 1), weight(bx + 1, by + 1, bz + 1), fx, fy, fz);
        dd = kf_interp3(dist(bx, by, bz), dist(bx + 1, by, bz), dist(bx, by + 1, bz), dist(bx + 1, by + 1, bz),
                        dist(bx, by, bz + 1), dist(bx + 1, by, bz + 1), dist(bx, by + 1, bz + 1),
                        dist(bx + 1, by + 1, bz + 1), fx, fy, fz);
    }
};

// SDF::update(displacementField), SDF.cpp:300-328.
template <class S>
__global__ void k_warp(Reader<S> R, const float *__restrict__ u, typename S::T *out_phi, int32_t *out_w)
{
    const GridDims g = R.g;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.n)
        return;
    int x, y, z;
    kf_xyz(g, i, x, y, z);
    const double lx = (double)x + 0.5 + (double)u[i];
    const double ly = (double)y + 0.5 + (double)u[g.n + i];
    const double lz = (double)z + 0.5 + (double)u[2 * g.n + i];
    double dd, ww;
    R.both(lx, ly, lz, dd, ww);
    out_phi[i] = S::enc((float)dd, R.scale);
    out_w[i] = (int32_t)(long long)ww;
}

// SDF::update for the listed voxels only (the voxels it can change; k_moved_mark): the new
// phi and weight of slot t, computed exactly as k_warp does, into compact arrays -- every
// read sees the unwarped SDF -- then written in place by k_warp_scatter.
template <class S>
__global__ void k_warp_list(Reader<S> R, const float *__restrict__ u, const unsigned int *__restrict__ list,
                            unsigned int count, typename S::T *__restrict__ new_phi, int32_t *__restrict__ new_w)
{
    const unsigned int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count)
        return;
    const GridDims g = R.g;
    const int64_t i = list[t];
    int x, y, z;
    kf_xyz(g, i, x, y, z);
    const double lx = (double)x + 0.5 + (double)u[i];
    const double ly = (double)y + 0.5 + (double)u[g.n + i];
    const double lz = (double)z + 0.5 + (double)u[2 * g.n + i];
    double dd, ww;
    R.both(lx, ly, lz, dd, ww);
    new_phi[t] = S::enc((float)dd, R.scale);
    new_w[t] = (int32_t)(long long)ww;
}

template <class S>
__global__ void k_warp_scatter(const unsigned int *__restrict__ list, unsigned int count,
                               const typename S::T *__restrict__ new_phi, const int32_t *__restrict__ new_w,
                               typename S::T *phi, int32_t *w)
{
    const unsigned int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= count)
        return;
    const unsigned int i = list[t];
    phi[i] = new_phi[t];
    w[i] = new_w[t];
}

// SDF::fuse(otherSdf), SDF.cpp:199-241: MERGE = FUSE_BY_MERGE (config.cpp:7), the weighted
// average; otherwise the value nearer the surface wins with its own weight.
// Four voxels per thread: the other volume's weights as one vector load, and a voxel only
// where the other volume has weight (most voxels of a live frame have none).
template <class S1, class S2, bool MERGE>
__global__ void k_fuse(typename S1::T *phi1, int32_t *w1, float scale1, double rscale1,
                       const typename S2::T *__restrict__ phi2, const int32_t *__restrict__ w2, float scale2,
                       int64_t n)
{
    const int64_t i0 = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) * 4;
    if (i0 >= n)
        return;
    int32_t bw[4];
    if (i0 + 4 <= n) {
        const int4 v = *reinterpret_cast<const int4 *>(w2 + i0);
        bw[0] = v.x;
        bw[1] = v.y;
        bw[2] = v.z;
        bw[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            bw[k] = i0 + k < n ? w2[i0 + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int32_t b = bw[k];
        if (b == 0)
            continue;
        const int64_t i = i0 + k;
        const int32_t a = w1[i];
        const double d2 = (double)S2::load(phi2, i, scale2);
        if (a == 0) {
            phi1[i] = S1::enc_r((float)d2, scale1, rscale1);
            w1[i] = b;
        } else if (MERGE) {
            const double d1 = (double)S1::load(phi1, i, scale1);
            const double nv = ((double)a * d1 + (double)b * d2) / (double)(a + b);
            phi1[i] = S1::enc_r((float)nv, scale1, rscale1);
            w1[i] = a + b;
        } else if (!(fabs((double)S1::load(phi1, i, scale1)) < fabs(d2))) {
            phi1[i] = S1::enc_r((float)d2, scale1, rscale1);
            w1[i] = b;
        }
    }
}

// SDF::fuse(otherSdf, displacementField), SDF.cpp:244-298.
template <class S1, class S2, bool MERGE>
__global__ void k_fuse_warped(typename S1::T *phi1, int32_t *w1, float scale1, Reader<S2> R,
                              const float *__restrict__ u)
{
    const GridDims g = R.g;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.n)
        return;
    int x, y, z;
    kf_xyz(g, i, x, y, z);
    const double lx = (double)x + 0.5 + (double)u[i];
    const double ly = (double)y + 0.5 + (double)u[g.n + i];
    const double lz = (double)z + 0.5 + (double)u[2 * g.n + i];
    double d2, b;
    R.both(lx, ly, lz, d2, b);
    if (fabs(b) < (double)1e-5f)
        return;
    const int32_t a = w1[i];
    if (a == 0) {
        phi1[i] = S1::enc((float)d2, scale1);
        w1[i] = (int32_t)(long long)b;
    } else if (MERGE) {
        const double d1 = (double)S1::load(phi1, i, scale1);
        const double nv = ((double)a * d1 + b * d2) / ((double)a + b);
        phi1[i] = S1::enc((float)nv, scale1);
        w1[i] = (int32_t)(long long)((double)a + b);
    } else if (!(fabs((double)S1::load(phi1, i, scale1)) < fabs(d2))) {
        phi1[i] = S1::enc((float)d2, scale1);
        w1[i] = (int32_t)(long long)b;
    }
}

__global__ void k_field_fill(float *u, int64_t n, float a, float b, float c)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) {
        u[i] = a;
        u[n + i] = b;
        u[2 * n + i] = c;
    }
}

__global__ void k_field_add(float *dst, const float *__restrict__ src, int64_t n3)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n3)
        dst[i] = dst[i] + src[i];
}

template <class S>
Reader<S> make_reader(const kf_volume *v)
{
    Reader<S> R;
    R.phi = (const typename S::T *)v->phi;
    R.w = v->w;
    R.g = grid_of(v);
    R.scale = v->scale();
    R.oob = v->oob();
    return R;
}

}  // namespace

kf_status kf_launch_fill_volume(kf_volume *v)
{
    kf_touch(v);
    KF_DISPATCH_DTYPE(v->dtype, S, {
        hipLaunchKernelGGL((k_fill<S>), dim3(blocks_for((v->n + 3) / 4)), dim3(BLOCK), 0, v->ctx->stream,
                           (typename S::T *)v->phi, v->w, v->n, S::enc((float)v->oob(), v->scale()));
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_upload_volume(kf_volume *v, const float *dphi, const int32_t *dw)
{
    kf_touch(v);
    (void)dw;
    KF_DISPATCH_DTYPE(v->dtype, S, {
        hipLaunchKernelGGL((k_encode<S>), dim3(blocks_for(v->n)), dim3(BLOCK), 0, v->ctx->stream,
                           (typename S::T *)v->phi, dphi, v->n, v->scale());
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_download_volume(const kf_volume *v, float *dphi)
{
    KF_DISPATCH_DTYPE(v->dtype, S, {
        hipLaunchKernelGGL((k_decode<S>), dim3(blocks_for(v->n)), dim3(BLOCK), 0, v->ctx->stream, dphi,
                           (const typename S::T *)v->phi, v->n, v->scale());
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_integrate(kf_volume *v, const float *ddepth, int w, int h, const double K[9],
                              const double W2C[16], double min_depth, double max_depth, bool fresh)
{
    kf_touch(v);
    KfCamera cam;
    for (int i = 0; i < 9; i++)
        cam.K[i] = K[i];
    for (int i = 0; i < 16; i++)
        cam.M[i] = W2C[i];
    cam.W = w;
    cam.H = h;
    double xb[3];
    for (int a = 0; a < 3; a++)
        xb[a] = kf_coord_bound(v->min_corner[a], v->voxel, v->dims[a]);
    kf_camera_prepare(cam, xb);
    const unsigned int nb = (unsigned int)((v->n + (int64_t)BLOCK * IPT - 1) / ((int64_t)BLOCK * IPT));
    KF_DISPATCH_DTYPE(v->dtype, S, {
        IntegrateArgs<S> A;
        A.phi = (typename S::T *)v->phi;
        A.w = v->w;
        A.g = grid_of(v);
        for (int a = 0; a < 3; a++)
            A.mn[a] = v->min_corner[a];
        A.voxel = v->voxel;
        A.clip = v->clip;
        A.scale = v->scale();
        A.rscale = kf_quot_rscale(A.scale);
        A.init = S::enc((float)v->oob(), A.scale);
        A.depth = ddepth;
        A.C = cam;
        A.min_depth = min_depth;
        A.max_depth = max_depth;
        const bool rowwise = kf_camera_rowwise(cam);
        const bool aligned = v->dims[0] % IPT == 0;
        auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, dim3(nb), dim3(BLOCK), 0, v->ctx->stream, A); };
        if (aligned) {
            if (fresh && rowwise)
                go(k_integrate<S, true, true, true>);
            else if (fresh)
                go(k_integrate<S, true, false, true>);
            else if (rowwise)
                go(k_integrate<S, false, true, true>);
            else
                go(k_integrate<S, false, false, true>);
        } else {
            if (fresh && rowwise)
                go(k_integrate<S, true, true, false>);
            else if (fresh)
                go(k_integrate<S, true, false, false>);
            else if (rowwise)
                go(k_integrate<S, false, true, false>);
            else
                go(k_integrate<S, false, false, false>);
        }
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_warp(kf_volume *v, const kf_field *u, void *tmp_phi, int32_t *tmp_w)
{
    kf_touch(v);
    KF_DISPATCH_DTYPE(v->dtype, S, {
        hipLaunchKernelGGL((k_warp<S>), dim3(blocks_for(v->n)), dim3(BLOCK), 0, v->ctx->stream, make_reader<S>(v),
                           (const float *)u->u(), (typename S::T *)tmp_phi, tmp_w);
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_fuse(kf_volume *can, const kf_volume *live, bool merge)
{
    kf_touch(can);
    KF_DISPATCH_DTYPE(can->dtype, S1, {
        KF_DISPATCH_DTYPE(live->dtype, S2, {
            const unsigned int nb = blocks_for((can->n + 3) / 4);
            const double rs = kf_quot_rscale(can->scale());
            if (merge)
                hipLaunchKernelGGL((k_fuse<S1, S2, true>), dim3(nb), dim3(BLOCK), 0, can->ctx->stream,
                                   (typename S1::T *)can->phi, can->w, can->scale(), rs,
                                   (const typename S2::T *)live->phi, (const int32_t *)live->w, live->scale(), can->n);
            else
                hipLaunchKernelGGL((k_fuse<S1, S2, false>), dim3(nb), dim3(BLOCK), 0, can->ctx->stream,
                                   (typename S1::T *)can->phi, can->w, can->scale(), rs,
                                   (const typename S2::T *)live->phi, (const int32_t *)live->w, live->scale(), can->n);
        });
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_fuse_warped(kf_volume *can, const kf_volume *live, const kf_field *u, bool merge)
{
    kf_touch(can);
    KF_DISPATCH_DTYPE(can->dtype, S1, {
        KF_DISPATCH_DTYPE(live->dtype, S2, {
            if (merge)
                hipLaunchKernelGGL((k_fuse_warped<S1, S2, true>), dim3(blocks_for(can->n)), dim3(BLOCK), 0,
                                   can->ctx->stream, (typename S1::T *)can->phi, can->w, can->scale(),
                                   make_reader<S2>(live), (const float *)u->u());
            else
                hipLaunchKernelGGL((k_fuse_warped<S1, S2, false>), dim3(blocks_for(can->n)), dim3(BLOCK), 0,
                                   can->ctx->stream, (typename S1::T *)can->phi, can->w, can->scale(),
                                   make_reader<S2>(live), (const float *)u->u());
        });
    });
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_field_fill(kf_field *f, const double v[3])
{
    hipLaunchKernelGGL(k_field_fill, dim3(blocks_for(f->n)), dim3(BLOCK), 0, f->ctx->stream, f->u(), f->n,
                       (float)v[0], (float)v[1], (float)v[2]);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_launch_field_add(kf_field *dst, const kf_field *src)
{
    hipLaunchKernelGGL(k_field_add, dim3(blocks_for(3 * dst->n)), dim3(BLOCK), 0, dst->ctx->stream, dst->u(),
                       (const float *)src->u(), 3 * dst->n);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

namespace {

// Warps the count voxels of ctx->active.list in place (k_warp_list, then k_warp_scatter).
kf_status warp_listed(kf_volume *v, const kf_field *u, unsigned int count)
{
    kf_touch(v);
    kf_context *ctx = v->ctx;
    kf_active_ws &a = ctx->active;
    if (count > 0) {
        const size_t need = (size_t)count * (v->elem_size() + sizeof(int32_t));
        if (need > a.wtmp_cap) {
            if (a.wtmp) {
                KF_HIP(hipStreamSynchronize(ctx->stream));
                KF_HIP(hipFree(a.wtmp));
                a.wtmp = nullptr;
                a.wtmp_cap = 0;
            }
            KF_HIP(hipMalloc((void **)&a.wtmp, need));
            a.wtmp_cap = need;
        }
        int32_t *nw = reinterpret_cast<int32_t *>(a.wtmp);
        void *np = a.wtmp + (size_t)count * sizeof(int32_t);
        const unsigned int nb = (count + BLOCK - 1) / BLOCK;
        KF_DISPATCH_DTYPE(v->dtype, S, {
            hipLaunchKernelGGL((k_warp_list<S>), dim3(nb), dim3(BLOCK), 0, ctx->stream, make_reader<S>(v),
                               (const float *)u->u(), a.list, count, (typename S::T *)np, nw);
            hipLaunchKernelGGL((k_warp_scatter<S>), dim3(nb), dim3(BLOCK), 0, ctx->stream, a.list, count,
                               (const typename S::T *)np, (const int32_t *)nw, (typename S::T *)v->phi, v->w);
        });
        KF_LAUNCH_CHECK();
    }
    return KF_OK;
}

}  // namespace

kf_status kf_warp_sparse(kf_volume *v, const kf_field *u, bool *done)
{
    *done = false;
    if (!kf_active_supported(v))
        return KF_OK;
    unsigned int count = 0;
    bool nonfinite = false;
    kf_status s = kf_list_moved(v, u->u(), &count, &nonfinite);
    if (s != KF_OK || nonfinite || (int64_t)count * 8 > v->n)
        return s;
    s = warp_listed(v, u, count);
    *done = s == KF_OK;
    return s;
}

kf_status kf_warp_after_deform(kf_volume *v, const kf_field *u, bool *done)
{
    *done = false;
    kf_active_ws &a = v->ctx->active;
    if (a.warp_vol != v || a.warp_field != u)
        return KF_OK;
    const unsigned int count = a.warp_count;
    a.warp_vol = nullptr;
    const kf_status s = warp_listed(v, u, count);
    *done = s == KF_OK;
    return s;
}
