// kf_project.hpp -- the pixel a voxel centre projects to in SDF::integrateDepthFrame
// (SDF.cpp:158-175), decided exactly as the reference's FP64 arithmetic decides it, and a
// cheaper route to the same decision.
//
// The reference's pixel is col = roundf(px + 0.5), row = roundf(py + 0.5), where
// (px, py) = K (c0 / c2, c1 / c2, 1) and c = W2C (X, Y, Z, 1) in FP64, narrowed to float before
// roundf (SDF.cpp:167,171).  kf_pixel_exact does exactly that: two FP64 divisions per voxel.
//
// kf_pixel_fast computes s = px + 0.5 another way: A = (K W2C)_row (X, Y, Z, 1) with the
// products K W2C formed once on the host, times a reciprocal of c2 refined from a float
// reciprocal.  The result differs from the exact path's s by a tiny amount E, and
// roundf(float(s)) is the integer nearest s except within half a float ulp of a half-integer.
// So wherever s lies farther than tau(s) = 2^-22 |s| + 2^-25 from every half-integer, the
// nearest integer of the fast s IS the reference's pixel; elsewhere (ok = false) the caller
// takes kf_pixel_exact.  tau covers:
//   * the float narrowing and the FP64 rounding of px + 0.5: at most 2^-24 |s| + 2^-53 |s|;
//   * the reciprocal: a float reciprocal within 2^-22 relative (the hardware's is within one
//     float ulp of the rounded c2), one Newton step -> 2^-44 + 2^-53 relative;
//   * the rounding of both FP64 evaluations against the real value of K c / c2.  This part
//     grows as |c2| shrinks; kf_camera_prepare bounds it from the grid's coordinate range and
//     sets c2min so that it stays below 2^-27 for |s| <= 2^13 (and far below |s| beyond that).
//     Voxels with |c2| < c2min take the exact path.
// Beyond |s| = 2^13 the voxel is outside any image of fewer than 4096 columns and rows
// whichever path decides, so the fast path calls it outside without a margin test.
// tests/cpp/project_check.cpp checks both paths against each other on the host, with the float
// reciprocal pushed to the ends of its error range, at voxels placed next to pixel boundaries.
#pragma once

#include <cmath>
#include <cstdint>

struct KfCamera {
    double K[9];   // intrinsics, row-major
    double M[16];  // world-to-camera pose, row-major
    double gx[4];  // row 0 of K * W2C[0:3]: px * c2 = gx . (X, Y, Z, 1)
    double gy[4];  // row 1
    double c2min;  // the fast path decides only where |c2| >= c2min (+inf: never)
    int W, H;
};

// The row constants of the exact projection for row (Y, Z): P[r] = (W2C[r][1] Y, W2C[r][2] Z,
// W2C[r][3] 1), as the reference's left-to-right sums form them.
__device__ __host__ inline void kf_row_exact(const KfCamera &C, double Y, double Z, double P[3][3])
{
    for (int r = 0; r < 3; r++) {
        P[r][0] = C.M[r * 4 + 1] * Y;
        P[r][1] = C.M[r * 4 + 2] * Z;
        P[r][2] = C.M[r * 4 + 3] * 1.0;
    }
}

// Camera-frame coordinate r of the voxel at X: ((W2C[r][0] X + W2C[r][1] Y) + W2C[r][2] Z) + W2C[r][3].
__device__ __host__ inline double kf_cam(const KfCamera &C, const double Pr[3], int r, double X)
{
    return ((C.M[r * 4 + 0] * X + Pr[0]) + Pr[1]) + Pr[2];
}

// The reference's pixel index row * W + col for the voxel at X of row P, or -1 outside the
// image (SDF.cpp:158-175).  c2 = kf_cam(C, P[2], 2, X).
__device__ __host__ inline int kf_pixel_exact(const KfCamera &C, const double P[3][3], double X, double c2)
{
    const double c0 = kf_cam(C, P[0], 0, X), c1 = kf_cam(C, P[1], 1, X);
    const double q0 = c0 / c2, q1 = c1 / c2;
    const double px = C.K[0] * q0 + C.K[1] * q1 + C.K[2] * 1.0;
    const double py = C.K[3] * q0 + C.K[4] * q1 + C.K[5] * 1.0;
    const float fc = roundf((float)(px + 0.5)), fr = roundf((float)(py + 0.5));
    // a coordinate that is NaN or beyond the int range lies outside the image (as the x86
    // conversion the reference compiles to makes it: INT_MIN)
    if (!(fabsf(fc) < 0x1p31f) || !(fabsf(fr) < 0x1p31f))
        return -1;
    const int col = (int)fc, row = (int)fr;
    return (col >= 0 && col < C.W && row >= 0 && row < C.H) ? row * C.W + col : -1;
}

// Row constants of the fast path: b = g1 Y + g2 Z + g3 for both image axes.
__device__ __host__ inline void kf_row_fast(const KfCamera &C, double Y, double Z, double &bx, double &by)
{
    bx = fma(C.gx[1], Y, fma(C.gx[2], Z, C.gx[3]));
    by = fma(C.gy[1], Y, fma(C.gy[2], Z, C.gy[3]));
}

// True when s = px + 0.5 decides its nearest integer beyond doubt (see the file comment).
__device__ __host__ inline bool kf_clear_of_half(double s)
{
    const double a = fabs(s);
    if (a >= 8192.0)
        return true;  // outside the image on either path (NaN fails both tests)
    const double fr = s - floor(s);
    return fabs(fr - 0.5) > fma(a, 0x1p-22, 0x1p-25);
}

// The pixel index (or -1) from the fast path, with ok = false where the exact path must decide.
// r0 is a float reciprocal of (float)c2 within 2^-22 relative.
__device__ __host__ inline int kf_pixel_fast(const KfCamera &C, double bx, double by, double X, double c2, float r0,
                                             bool &ok)
{
    const double r1 = (double)r0;
    const double r = fma(r1, fma(-c2, r1, 1.0), r1);
    const double sx = fma(fma(C.gx[0], X, bx), r, 0.5);
    const double sy = fma(fma(C.gy[0], X, by), r, 0.5);
    ok = fabs(c2) >= C.c2min && kf_clear_of_half(sx) && kf_clear_of_half(sy);
    const bool in = sx > -0.5 && sx < (double)C.W - 0.5 && sy > -0.5 && sy < (double)C.H - 0.5;
    return in ? (int)rint(sy) * C.W + (int)rint(sx) : -1;
}

// Cameras whose depth and image row stay fixed along a grid row: W2C[2][0] = 0 and gy[0] = 0
// (the reference's per-frame identity pose, KillingFusion.cpp:226, among them).  Then
// c2 = ((W2C[2][0] X + P0) + P1) + P2 is the same double for every X of the row whenever it is
// nonzero (W2C[2][0] X is a zero, and adding a zero leaves a nonzero sum unchanged), and the
// fast path's s_y is the same for every voxel of the row.  The row's s_x is x a + b, a linear
// form in the column index x whose roundings the launch's coordinate bound covers (it bounds
// |mn_x| + voxel (n_x + 1), so the products and sums below stay within the 16u charged to A).
struct KfRowwise {
    double c2;     // camera depth of every voxel of the row (valid where ok)
    double a, b;   // s_x of column x = x a + b
    int rowbase;   // image row * W, or -1 when the row projects outside the image
    bool ok;       // the row's own decisions (|c2| >= c2min, the image row) are clear
};

__device__ __host__ inline bool kf_camera_rowwise(const KfCamera &C)
{
    return C.M[8] == 0.0 && C.gy[0] == 0.0 && C.c2min < INFINITY;
}

// The row's camera depth: the value kf_cam gives at any X of the row (where nonzero).
__device__ __host__ inline double kf_rowwise_depth(const KfCamera &C, double Y, double Z)
{
    return ((0.0 + C.M[9] * Y) + C.M[10] * Z) + C.M[11] * 1.0;
}

// The rest of the row set-up, with r0 a float reciprocal of (float)c2 within 2^-22 relative;
// the column x of the row sits at X = voxel (x + 0.5) + mnx.
__device__ __host__ inline KfRowwise kf_rowwise_row(const KfCamera &C, double Y, double Z, double c2, float r0,
                                                    double voxel, double mnx)
{
    KfRowwise R;
    R.c2 = c2;
    double bx, by;
    kf_row_fast(C, Y, Z, bx, by);
    const double r1 = (double)r0;
    const double r = fma(r1, fma(-c2, r1, 1.0), r1);
    const double sy = fma(by, r, 0.5);
    R.ok = fabs(c2) >= C.c2min && kf_clear_of_half(sy);
    R.rowbase = (sy > -0.5 && sy < (double)C.H - 0.5) ? (int)rint(sy) * C.W : -1;
    R.a = (C.gx[0] * voxel) * r;
    R.b = fma(fma(C.gx[0], fma(0.5, voxel, mnx), bx), r, 0.5);
    return R;
}

// The pixel index (or -1) of column x of a rowwise row, with ok = false where the exact path
// must decide.
__device__ __host__ inline int kf_pixel_rowwise(const KfCamera &C, const KfRowwise &R, int x, bool &ok)
{
    const double sx = fma((double)x, R.a, R.b);
    ok = R.ok && kf_clear_of_half(sx);
    const bool in = R.rowbase >= 0 && sx > -0.5 && sx < (double)C.W - 0.5;
    return in ? R.rowbase + (int)rint(sx) : -1;
}

// Bound on |X|, |Y|, |Z| (and on |mn| + |x voxel| for the rowwise form) over a grid of n voxels
// of size voxel from corner mn, per axis.
inline double kf_coord_bound(double mn, double voxel, int n)
{
    return std::fabs(mn) + std::fabs(voxel) * (n + 1.0);
}

// Host set-up: the products K W2C and the threshold c2min for a grid whose voxel centres have
// |X| <= xb[0], |Y| <= xb[1], |Z| <= xb[2].  The bound on the gap between the two paths' s
// (both against the real value A / c2):
//   exact path:  4u T + u (|K0 q0| + |K1 q1|) + (|K0| (d0 + |q0| d2) + |K1| (d1 + |q1| d2)) / |c2|,
//                T = |K0 q0| + |K1 q1| + |K2| (the sums of px and the two quotients' roundings)
//   fast path:   u |s| + 2^-43 |s| + (dA + |px| d2) / |c2|
// with u = 2^-53, d_r >= the rounding error of c_r (4 operations on terms of total size
// Cmag_r, charged 8u Cmag_r), dA >= that of A (host products, two row and one voxel
// multiply-add, charged 16u Amag), |q_r| <= Cmag_r / |c2| and |px| <= 2^14.  Collected as
// a1 / |c2| + a2 / c2^2 + b (every coefficient charged at least its first-order size above);
// c2min makes each of the first two <= 2^-28, and b must be <= 2^-28.
inline void kf_camera_prepare(KfCamera &C, const double xb[3])
{
    const double u = 0x1p-53;
    const double vb[4] = {xb[0], xb[1], xb[2], 1.0};
    double cmag[3];
    bool finite = true;
    for (int i = 0; i < 9; i++)
        finite = finite && std::isfinite(C.K[i]);
    for (int i = 0; i < 16; i++)
        finite = finite && std::isfinite(C.M[i]);
    for (int j = 0; j < 3; j++)
        finite = finite && std::isfinite(xb[j]);
    for (int r = 0; r < 3; r++) {
        cmag[r] = 0.0;
        for (int j = 0; j < 4; j++)
            cmag[r] += std::fabs(C.M[r * 4 + j]) * vb[j];
    }
    double c2min = 0x1p-100;
    double b = 0.0;
    for (int i = 0; i < 2; i++) {
        const double *k = C.K + 3 * i;
        double *g = i == 0 ? C.gx : C.gy;
        double amag = 0.0;
        for (int j = 0; j < 4; j++) {
            g[j] = k[0] * C.M[j] + k[1] * C.M[4 + j] + k[2] * C.M[8 + j];
            amag += (std::fabs(k[0] * C.M[j]) + std::fabs(k[1] * C.M[4 + j]) + std::fabs(k[2] * C.M[8 + j])) * vb[j];
        }
        const double d0 = 8 * u * cmag[0], d1 = 8 * u * cmag[1], d2 = 8 * u * cmag[2], dA = 16 * u * amag;
        const double kc = std::fabs(k[0]) * cmag[0] + std::fabs(k[1]) * cmag[1];
        const double a1 = 8 * u * kc + std::fabs(k[0]) * d0 + std::fabs(k[1]) * d1 + dA + 0x1p14 * d2;
        const double a2 = kc * d2;
        c2min = std::fmax(c2min, std::fmax(a1 * 0x1p28, std::sqrt(a2 * 0x1p28)));
        b = std::fmax(b, 8 * u * std::fabs(k[2]) + 4 * u * 0x1p14);
    }
    const bool usable = finite && C.W > 0 && C.H > 0 && C.W < 4096 && C.H < 4096 && b <= 0x1p-28 &&
                        cmag[2] < 0x1p100 && std::isfinite(c2min);
    C.c2min = usable ? c2min : INFINITY;
}
