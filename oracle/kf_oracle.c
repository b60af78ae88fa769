/*
 * kf_oracle.c -- FP64 CPU restatement of KillingFusion's per-frame hot path.
 * TEST INFRASTRUCTURE ONLY (see kf_oracle.h): the checker and the timed CPU
 * baseline, never part of the shipped path.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Contraction is
 * off so every product and sum rounds separately, as the reference's
 * x86-64 build (no -march, hence no FMA) does.
 */
#include "kf_oracle.h"
#include "kf_mc_table.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(g, x, y, z) ((int64_t)(z) * (g)->n[0] * (g)->n[1] + (int64_t)(y) * (g)->n[0] + (x))
#define NVOX(g) ((int64_t)(g)->n[0] * (g)->n[1] * (g)->n[2])

/* ------------------------------------------------------------------------ */
/* a1: grid sizing and allocation                                           */
/* ------------------------------------------------------------------------ */

/* SDF::computeVoxelGridSize, SDF.cpp:117-124: int(bound/voxel)+1 per axis. */
void kfo_dims_from_bounds(const double min[3], const double max[3], double voxel, int32_t out[3])
{
    for (int a = 0; a < 3; a++)
        out[a] = (int32_t)((max[a] - min[a]) / voxel) + 1;
}

/* SDF::allocateMemoryForSDF, SDF.cpp:126-135: phi = Max+eps, w = 0. */
void kfo_sdf_init(double *phi, int64_t *w, int64_t n, double oob)
{
    for (int64_t i = 0; i < n; i++) {
        phi[i] = oob;
        w[i] = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* a3: interpolation chain x -> y -> z (utils.h:6-46)                       */
/* ------------------------------------------------------------------------ */

static inline double lerp1(double v0, double v1, double x) { return v0 * (1 - x) + v1 * x; }

double kfo_interp3(double v000, double v001, double v010, double v011,
                   double v100, double v101, double v110, double v111,
                   double x, double y, double z)
{
    double s = lerp1(lerp1(v000, v001, x), lerp1(v010, v011, x), y);
    double t = lerp1(lerp1(v100, v101, x), lerp1(v110, v111, x), y);
    return lerp1(s, t, z);
}

/* ------------------------------------------------------------------------ */
/* a4: bounds-checked node reads (SDF.cpp:462-501)                           */
/* ------------------------------------------------------------------------ */

static inline int in_grid(const kfo_grid *g, int x, int y, int z)
{
    return x >= 0 && x < g->n[0] && y >= 0 && y < g->n[1] && z >= 0 && z < g->n[2];
}

/* SDF::getDistanceAtIndex, SDF.cpp:482-487: out of grid -> Max+eps. */
double kfo_distance_at_index(const double *phi, const kfo_grid *g, int x, int y, int z)
{
    if (!in_grid(g, x, y, z))
        return g->oob;
    return phi[IDX(g, x, y, z)];
}

/* SDF::getWeightAtIndex, SDF.cpp:496-501: out of grid -> 0. */
int64_t kfo_weight_at_index(const int64_t *w, const kfo_grid *g, int x, int y, int z)
{
    if (!in_grid(g, x, y, z))
        return 0;
    return w[IDX(g, x, y, z)];
}

/* ------------------------------------------------------------------------ */
/* a5: cell-centred trilinear reads (SDF.cpp:503-524, 584-604)               */
/* ------------------------------------------------------------------------ */

/* SDF::getDistance(Vector3d), SDF.cpp:503-524.  loc - 0.5, base = int cast
 * (truncation toward zero), weights = loc' - base. */
double kfo_get_distance(const double *phi, const kfo_grid *g, const double loc[3])
{
    double t[3];
    int b[3];
    for (int a = 0; a < 3; a++) {
        t[a] = loc[a] - 0.5;
        b[a] = (int)t[a];
    }
    double v000 = kfo_distance_at_index(phi, g, b[0], b[1], b[2]);
    double v001 = kfo_distance_at_index(phi, g, b[0] + 1, b[1], b[2]);
    double v010 = kfo_distance_at_index(phi, g, b[0], b[1] + 1, b[2]);
    double v011 = kfo_distance_at_index(phi, g, b[0] + 1, b[1] + 1, b[2]);
    double v100 = kfo_distance_at_index(phi, g, b[0], b[1], b[2] + 1);
    double v101 = kfo_distance_at_index(phi, g, b[0] + 1, b[1], b[2] + 1);
    double v110 = kfo_distance_at_index(phi, g, b[0], b[1] + 1, b[2] + 1);
    double v111 = kfo_distance_at_index(phi, g, b[0] + 1, b[1] + 1, b[2] + 1);
    return kfo_interp3(v000, v001, v010, v011, v100, v101, v110, v111,
                       t[0] - (double)b[0], t[1] - (double)b[1], t[2] - (double)b[2]);
}

/* SDF::getWeight(Vector3d), SDF.cpp:584-604 (integer node weights read as double). */
double kfo_get_weight(const int64_t *w, const kfo_grid *g, const double loc[3])
{
    double t[3];
    int b[3];
    for (int a = 0; a < 3; a++) {
        t[a] = loc[a] - 0.5;
        b[a] = (int)t[a];
    }
    double v000 = (double)kfo_weight_at_index(w, g, b[0], b[1], b[2]);
    double v001 = (double)kfo_weight_at_index(w, g, b[0] + 1, b[1], b[2]);
    double v010 = (double)kfo_weight_at_index(w, g, b[0], b[1] + 1, b[2]);
    double v011 = (double)kfo_weight_at_index(w, g, b[0] + 1, b[1] + 1, b[2]);
    double v100 = (double)kfo_weight_at_index(w, g, b[0], b[1], b[2] + 1);
    double v101 = (double)kfo_weight_at_index(w, g, b[0] + 1, b[1], b[2] + 1);
    double v110 = (double)kfo_weight_at_index(w, g, b[0], b[1] + 1, b[2] + 1);
    double v111 = (double)kfo_weight_at_index(w, g, b[0] + 1, b[1] + 1, b[2] + 1);
    return kfo_interp3(v000, v001, v010, v011, v100, v101, v110, v111,
                       t[0] - (double)b[0], t[1] - (double)b[1], t[2] - (double)b[2]);
}

/* ------------------------------------------------------------------------ */
/* a7: displacement-field reads (DisplacementField.cpp:22-63)               */
/* ------------------------------------------------------------------------ */

/* DisplacementField::getDisplacementAt, DisplacementField.cpp:30-38.  Only the
 * linear index is range-checked, so x = nx (or y = ny) reads the first voxel of
 * the next row (slice). */
void kfo_disp_at(const double *u, const kfo_grid *g, int x, int y, int z, double out[3])
{
    int64_t n = NVOX(g);
    int64_t i = (int64_t)z * g->n[0] * g->n[1] + (int64_t)y * g->n[0] + x;
    if (i < 0 || i >= n) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    out[0] = u[i];
    out[1] = u[n + i];
    out[2] = u[2 * n + i];
}

/* DisplacementField::getDisplacementAtf, DisplacementField.cpp:40-58: no -0.5
 * offset, truncating cast, lerp chain of utils.h:29-46 per component. */
void kfo_disp_at_f(const double *u, const kfo_grid *g, const double loc[3], double out[3])
{
    int b[3];
    for (int a = 0; a < 3; a++)
        b[a] = (int)loc[a];
    double c[8][3];
    kfo_disp_at(u, g, b[0], b[1], b[2], c[0]);
    kfo_disp_at(u, g, b[0] + 1, b[1], b[2], c[1]);
    kfo_disp_at(u, g, b[0], b[1] + 1, b[2], c[2]);
    kfo_disp_at(u, g, b[0] + 1, b[1] + 1, b[2], c[3]);
    kfo_disp_at(u, g, b[0], b[1], b[2] + 1, c[4]);
    kfo_disp_at(u, g, b[0] + 1, b[1], b[2] + 1, c[5]);
    kfo_disp_at(u, g, b[0], b[1] + 1, b[2] + 1, c[6]);
    kfo_disp_at(u, g, b[0] + 1, b[1] + 1, b[2] + 1, c[7]);
    double fx = loc[0] - (double)b[0], fy = loc[1] - (double)b[1], fz = loc[2] - (double)b[2];
    for (int k = 0; k < 3; k++)
        out[k] = kfo_interp3(c[0][k], c[1][k], c[2][k], c[3][k], c[4][k], c[5][k], c[6][k], c[7][k],
                             fx, fy, fz);
}

/* ------------------------------------------------------------------------ */
/* a6: warped reads (SDF.cpp:570-582)                                        */
/* ------------------------------------------------------------------------ */

/* SDF::getDistance(idx, disp), SDF.cpp:570-575: phi at idx + u[idx] (u read at
 * the integer index, phi trilinear). */
double kfo_distance_idx_disp(const double *phi, const double *u, const kfo_grid *g, int x, int y, int z)
{
    double d[3];
    kfo_disp_at(u, g, x, y, z, d);
    double loc[3] = {(double)x + d[0] + 0.5, (double)y + d[1] + 0.5, (double)z + d[2] + 0.5};
    return kfo_get_distance(phi, g, loc);
}

/* SDF::getDistancef(loc, disp), SDF.cpp:577-582: loc + u_f(loc) + 0.5. */
double kfo_distance_f(const double *phi, const double *u, const kfo_grid *g, const double loc[3])
{
    double d[3];
    kfo_disp_at_f(u, g, loc, d);
    double p[3] = {loc[0] + d[0] + 0.5, loc[1] + d[1] + 0.5, loc[2] + d[2] + 0.5};
    return kfo_get_distance(phi, g, p);
}

/* ------------------------------------------------------------------------ */
/* a8: displaced central-difference gradient (SDF.cpp:726-744)               */
/* ------------------------------------------------------------------------ */

void kfo_distance_gradient_disp(const double *phi, const double *u, const kfo_grid *g,
                                int x, int y, int z, double h, double out[3])
{
    double gl[3] = {(double)x + 0.5, (double)y + 0.5, (double)z + 0.5};
    for (int a = 0; a < 3; a++) {
        double lp[3] = {gl[0], gl[1], gl[2]};
        double lm[3] = {gl[0], gl[1], gl[2]};
        lp[a] = gl[a] + h;
        lm[a] = gl[a] - h;
        out[a] = kfo_distance_f(phi, u, g, lp) - kfo_distance_f(phi, u, g, lm);
    }
    for (int a = 0; a < 3; a++)
        out[a] = out[a] / (2 * h);
}

/* ------------------------------------------------------------------------ */
/* a9: displaced 19-sample Hessian (SDF.cpp:842-886)                         */
/* ------------------------------------------------------------------------ */

/* Sample offsets in units of h, in the order of SDF.cpp:847-865. */
static const int HESS_OFF[19][3] = {
    {0, 0, 0},                                        /* fxyz */
    {2, 0, 0}, {-2, 0, 0}, {0, 2, 0}, {0, -2, 0},     /* fxx, fyy */
    {0, 0, 2}, {0, 0, -2},                            /* fzz */
    {1, 0, 1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, 1},     /* +z diagonals */
    {1, 1, 0}, {-1, 1, 0},                            /* +y diagonals */
    {-1, 0, -1}, {1, 0, -1}, {0, -1, -1}, {0, 1, -1}, /* -z diagonals */
    {-1, -1, 0}, {1, -1, 0}};                         /* -y diagonals */

/* Combine 19 samples f[k] (HESS_OFF order) into the symmetric 3x3 Hessian,
 * SDF.cpp:867-885 (row-major out[9]). */
static void hessian_combine(const double f[19], double h, double out[9])
{
    double fxyz = f[0];
    double fxx = (f[1] - 2 * fxyz + f[2]);
    double fyy = (f[3] - 2 * fxyz + f[4]);
    double fzz = (f[5] - 2 * fxyz + f[6]);
    /* fxz = (x+1,z+1) + (x-1,z-1) - (x+1,z-1) - (x-1,z+1) */
    double fxz = (f[7] + f[13] - f[14] - f[8]);
    /* fyz = (y+1,z+1) + (y-1,z-1) - (y+1,z-1) - (y-1,z+1) */
    double fyz = (f[9] + f[15] - f[16] - f[10]);
    /* fxy = (x+1,y+1) + (x-1,y-1) - (x+1,y-1) - (x-1,y+1) */
    double fxy = (f[11] + f[17] - f[18] - f[12]);
    double den = (4 * h * h);
    out[0] = fxx / den; out[1] = fxy / den; out[2] = fxz / den;
    out[3] = fxy / den; out[4] = fyy / den; out[5] = fyz / den;
    out[6] = fxz / den; out[7] = fyz / den; out[8] = fzz / den;
}

void kfo_distance_hessian_disp(const double *phi, const double *u, const kfo_grid *g,
                               int x, int y, int z, double h, double out[9])
{
    double gl[3] = {(double)x + 0.5, (double)y + 0.5, (double)z + 0.5};
    double f[19];
    for (int k = 0; k < 19; k++) {
        /* gridLocation + deltaSize * Vector3d(o) */
        double loc[3];
        for (int a = 0; a < 3; a++)
            loc[a] = gl[a] + h * (double)HESS_OFF[k][a];
        f[k] = kfo_distance_f(phi, u, g, loc);
    }
    hessian_combine(f, h, out);
}

/* ------------------------------------------------------------------------ */
/* a12: non-displaced stencils, used by the reference self-tests             */
/* ------------------------------------------------------------------------ */

/* SDF::computeDistanceGradient(Vector3d), SDF.cpp:657-679: unit step, returns
 * 0 (and reports 0) when the location is outside [0, n). */
int kfo_distance_gradient_plain(const double *phi, const kfo_grid *g, const double loc[3], double out[3])
{
    for (int a = 0; a < 3; a++) {
        if (loc[a] < 0.0 || loc[a] >= (double)g->n[a]) {
            out[0] = out[1] = out[2] = 0.0;
            return 0;
        }
    }
    for (int a = 0; a < 3; a++) {
        double lp[3] = {loc[0], loc[1], loc[2]};
        double lm[3] = {loc[0], loc[1], loc[2]};
        lp[a] += 1.0;
        lm[a] -= 1.0;
        out[a] = (kfo_get_distance(phi, g, lp) - kfo_get_distance(phi, g, lm)) / 2;
    }
    return 1;
}

/* SDF::computeDistanceHessian(Vector3d), SDF.cpp:746-795 (the range test is on
 * loc - 0.5). */
int kfo_distance_hessian_plain(const double *phi, const kfo_grid *g, const double loc[3], double h, double out[9])
{
    for (int a = 0; a < 3; a++) {
        double t = loc[a] - 0.5;
        if (t < 0.0 || t >= (double)g->n[a]) {
            for (int k = 0; k < 9; k++)
                out[k] = 0.0;
            return 0;
        }
    }
    double f[19];
    for (int k = 0; k < 19; k++) {
        double l[3];
        for (int a = 0; a < 3; a++)
            l[a] = loc[a] + h * (double)HESS_OFF[k][a];
        f[k] = kfo_get_distance(phi, g, l);
    }
    hessian_combine(f, h, out);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* a11: Jacobian and Killing energy density (DisplacementField.cpp:90-193)   */
/* ------------------------------------------------------------------------ */

/* DisplacementField::computeJacobian, :90-140.  out row-major, column i = d/dx_i. */
void kfo_jacobian(const double *u, const kfo_grid *g, double x, double y, double z, double h, double out[9])
{
    for (int i = 0; i < 3; i++) {
        double lp[3] = {x, y, z}, lm[3] = {x, y, z};
        lp[i] += h;
        lm[i] -= h;
        double dp[3], dm[3];
        kfo_disp_at_f(u, g, lp, dp);
        kfo_disp_at_f(u, g, lm, dm);
        for (int r = 0; r < 3; r++)
            out[r * 3 + i] = (dp[r] - dm[r]) / (2 * h);
    }
}

/* DisplacementField::computeKillingEnergy, :171-193: ||J||_F^2 + gamma tr(J^T J...)
 * as vec(J).vec(J) + gamma vec(J^T).vec(J); non-finite J gives 0. */
double kfo_killing_energy(const double *u, const kfo_grid *g, double x, double y, double z,
                          double h, double gamma)
{
    double J[9];
    kfo_jacobian(u, g, x, y, z, h, J);
    for (int k = 0; k < 9; k++)
        if (!isfinite(J[k]))
            return 0.0;
    double a = 0.0, b = 0.0;
    /* Eigen maps are column-major: vec index = c*3 + r. */
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
            double jv = J[r * 3 + c];
            double jt = J[c * 3 + r]; /* (J^T)(r,c) */
            a += jv * jv;
            b += jt * jv;
        }
    return a + gamma * b;
}

/* ------------------------------------------------------------------------ */
/* a10: variational Killing gradient with clamp (DisplacementField.cpp:301-347) */
/* ------------------------------------------------------------------------ */

void kfo_killing_gradient2(const double *u, const kfo_grid *g, int x, int y, int z,
                           double h, double gamma, double clamp, double out[3])
{
    double gl[3] = {(double)x, (double)y, (double)z};
    double d[19][3];
    for (int k = 0; k < 19; k++) {
        double loc[3];
        for (int a = 0; a < 3; a++)
            loc[a] = gl[a] + h * (double)HESS_OFF[k][a];
        kfo_disp_at_f(u, g, loc, d[k]);
    }
    double dxx[3], dyy[3], dzz[3], dxy[3], dxz[3], dyz[3];
    for (int c = 0; c < 3; c++) {
        dxx[c] = (d[1][c] - 2 * d[0][c] + d[2][c]);
        dyy[c] = (d[3][c] - 2 * d[0][c] + d[4][c]);
        dzz[c] = (d[5][c] - 2 * d[0][c] + d[6][c]);
        dxz[c] = (d[7][c] + d[13][c] - d[14][c] - d[8][c]);
        dyz[c] = (d[9][c] + d[15][c] - d[16][c] - d[10][c]);
        dxy[c] = (d[11][c] + d[17][c] - d[18][c] - d[12][c]);
    }
    double div[3] = {dxx[0] + dxy[1] + dxz[2],
                     dxy[0] + dyy[1] + dyz[2],
                     dxz[0] + dyz[1] + dzz[2]};
    double kg[3];
    for (int c = 0; c < 3; c++)
        kg[c] = -2 * (dxx[c] + dyy[c] + dzz[c]) + -2 * gamma * div[c];
    double den = (4 * h * h);
    double nrm = sqrt(kg[0] * kg[0] + kg[1] * kg[1] + kg[2] * kg[2]);
    if (nrm > clamp) {
        for (int c = 0; c < 3; c++)
            kg[c] = kg[c] / nrm * clamp;
    }
    for (int c = 0; c < 3; c++)
        out[c] = kg[c] / den;
}

/* ------------------------------------------------------------------------ */
/* a13-a16: energy gradient (KillingFusion.cpp:399-476)                      */
/* ------------------------------------------------------------------------ */

/* computeDataEnergyGradient, KillingFusion.cpp:444-455. */
void kfo_data_gradient(const double *live, const double *can, const double *u, const kfo_grid *g,
                       const kfo_params *p, int x, int y, int z, double out[3])
{
    double gr[3];
    kfo_distance_gradient_disp(live, u, g, x, y, z, p->h, gr);
    double src = kfo_distance_idx_disp(live, u, g, x, y, z);
    double dst = kfo_distance_at_index(can, g, x, y, z);
    double s = (src - dst) / p->voxel_size;
    for (int a = 0; a < 3; a++)
        out[a] = s * gr[a];
}

/* computeLevelSetEnergyGradient, KillingFusion.cpp:464-476:
 * H g (|g| - 1) / (|g| + eps). */
void kfo_levelset_gradient(const double *live, const double *u, const kfo_grid *g,
                           const kfo_params *p, int x, int y, int z, double out[3])
{
    double gr[3], H[9];
    kfo_distance_gradient_disp(live, u, g, x, y, z, p->h, gr);
    kfo_distance_hessian_disp(live, u, g, x, y, z, p->h, H);
    double n = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    for (int r = 0; r < 3; r++) {
        double hg = H[r * 3 + 0] * gr[0] + H[r * 3 + 1] * gr[1] + H[r * 3 + 2] * gr[2];
        out[r] = hg * (n - 1) / (n + p->epsilon);
    }
}

/* computeEnergyGradient, KillingFusion.cpp:399-442:
 * data + omega_k * killing + omega_s * levelset, each switched by EnergyTypeUsed. */
void kfo_energy_gradient(const double *live, const double *can, const double *u, const kfo_grid *g,
                         const kfo_params *p, int x, int y, int z, double out[3])
{
    double dg[3] = {0, 0, 0}, lg[3] = {0, 0, 0}, kg[3] = {0, 0, 0};
    if (p->use_data)
        kfo_data_gradient(live, can, u, g, p, x, y, z, dg);
    if (p->use_levelset) {
        kfo_levelset_gradient(live, u, g, p, x, y, z, lg);
        for (int a = 0; a < 3; a++)
            lg[a] = lg[a] * p->omega_s;
    }
    if (p->use_killing) {
        kfo_killing_gradient2(u, g, x, y, z, p->h, p->gamma, p->killing_clamp, kg);
        for (int a = 0; a < 3; a++)
            kg[a] = kg[a] * p->omega_k;
    }
    for (int a = 0; a < 3; a++)
        out[a] = (dg[a] + kg[a]) + lg[a];
}

/* Skip test of the all-voxels loop, KillingFusion.cpp:272-274. */
static inline int jacobi_active(double d, const kfo_params *p)
{
    return !(d > p->max_surface - p->epsilon || d < -p->clip);
}

void kfo_energy_gradient_grid(const double *live, const double *can, const double *u,
                              const kfo_grid *g, const kfo_params *p, int masked, double *grad)
{
    int64_t n = NVOX(g);
#pragma omp parallel for schedule(dynamic)
    for (int z = 0; z < g->n[2]; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                int64_t i = IDX(g, x, y, z);
                double o[3] = {0, 0, 0};
                if (!masked || jacobi_active(kfo_distance_idx_disp(live, u, g, x, y, z), p))
                    kfo_energy_gradient(live, can, u, g, p, x, y, z, o);
                grad[i] = o[0];
                grad[n + i] = o[1];
                grad[2 * n + i] = o[2];
            }
}

/* ------------------------------------------------------------------------ */
/* a17: optimiser                                                           */
/* ------------------------------------------------------------------------ */

/* computeDisplacementField, all-voxels mode with UsePreviousIterationDeformationField
 * (KillingFusion.cpp:248-351): every iteration reads the frozen field, collects
 * -alpha*grad for active voxels into a zero field, then adds it (Jacobi).  The
 * stop rule (max |du| < stop_max_update) is not in the reference loop, which
 * always runs max_iters; it is off when stop_max_update <= 0.  Returns 0, or 1
 * when a non-finite displacement appears (st->diverged_index = lowest index). */
/* Trust region of the all-voxels mode, KillingFusion.cpp:288-314 (UseTrustStrategy with the data
 * energy alone): du = -alpha g is shrunk (alpha / 1.5 each time) until the data residual at the
 * moved sample, |src - dest|, is no larger than at the current one, |prev - dest|; src is read at
 * ((j + u_j) + du) + 0.5 through SDF::getDistance(Vector3d) (SDF.cpp:503-524).  Returns 0 when
 * alpha fell below 1e-7: the voxel takes no step and the reference's `break` leaves its x loop,
 * so the rest of the row takes none either in this iteration. */
static int trust_step(const double *live, const double *can, const double *u, const kfo_grid *g,
                      const kfo_params *p, int x, int y, int z, const double gr[3], double du[3])
{
    double a = p->alpha;
    const double dest = kfo_distance_at_index(can, g, x, y, z);
    const double prev = kfo_distance_idx_disp(live, u, g, x, y, z);
    double uj[3];
    kfo_disp_at(u, g, x, y, z, uj);
    int decreased = 0;
    do {
        const double loc[3] = {((double)x + uj[0] + du[0]) + 0.5, ((double)y + uj[1] + du[1]) + 0.5,
                               ((double)z + uj[2] + du[2]) + 0.5};
        const double src = kfo_get_distance(live, g, loc);
        const double conv = fabs(src - dest) - fabs(prev - dest);
        if (conv > 0) {
            a /= 1.5;
            for (int k = 0; k < 3; k++)
                du[k] = -a * gr[k];
        } else {
            decreased = 1;
        }
    } while (!decreased && a > 1e-7);
    return !(a < 1e-7);
}

int kfo_deform_jacobi(const double *live, const double *can, double *u, const kfo_grid *g,
                      const kfo_params *p, kfo_stats *st)
{
    int64_t n = NVOX(g);
    double *delta = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *zmax = (double *)malloc(sizeof(double) * (size_t)g->n[2]);
    int64_t *zact = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->n[2]);
    st->iters = 0;
    st->diverged_index = -1;
    st->max_update = 0.0;
    st->energy = 0.0;
    st->active = 0;
    const int trust = p->use_trust && p->use_data && !p->use_levelset && !p->use_killing;
    for (int it = 0; it < p->max_iters; it++) {
        memset(delta, 0, sizeof(double) * 3 * (size_t)n);
#pragma omp parallel for schedule(dynamic)
        for (int z = 0; z < g->n[2]; z++) {
            double m = 0.0;
            int64_t act = 0;
            for (int y = 0; y < g->n[1]; y++)
                for (int x = 0; x < g->n[0]; x++) {
                    double d = kfo_distance_idx_disp(live, u, g, x, y, z);
                    if (!jacobi_active(d, p))
                        continue;
                    act++;
                    double gr[3];
                    kfo_energy_gradient(live, can, u, g, p, x, y, z, gr);
                    int64_t i = IDX(g, x, y, z);
                    double du[3] = {-p->alpha * gr[0], -p->alpha * gr[1], -p->alpha * gr[2]};
                    if (trust && !trust_step(live, can, u, g, p, x, y, z, gr, du))
                        break;
                    delta[i] = du[0];
                    delta[n + i] = du[1];
                    delta[2 * n + i] = du[2];
                    double nn = sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
                    if (nn > m)
                        m = nn;
                }
            zmax[z] = m;
            zact[z] = act;
        }
        double m = 0.0;
        int64_t act = 0;
        for (int z = 0; z < g->n[2]; z++) {
            if (zmax[z] > m)
                m = zmax[z];
            act += zact[z];
        }
        for (int64_t i = 0; i < 3 * n; i++)
            u[i] = u[i] + delta[i];
        st->iters = it + 1;
        st->max_update = m;
        st->active = act;
        for (int64_t i = 0; i < n; i++) {
            if (!isfinite(u[i]) || !isfinite(u[n + i]) || !isfinite(u[2 * n + i])) {
                st->diverged_index = i;
                break;
            }
        }
        if (st->diverged_index >= 0)
            break;
        if (p->stop_max_update > 0.0 && m < p->stop_max_update)
            break;
    }
    free(delta);
    free(zmax);
    free(zact);
    return st->diverged_index >= 0;
}

/* computeDisplacementField, per-voxel mode (KillingFusion.cpp:353-396), run
 * serially in z->y->x order so it is deterministic (the reference runs it under
 * an OpenMP team with in-place writes, which races).  Skip test is symmetric
 * +-(Max - eps) (:369). */
int kfo_deform_gs_serial(const double *live, const double *can, double *u, const kfo_grid *g,
                         const kfo_params *p, kfo_stats *st)
{
    int64_t n = NVOX(g);
    st->iters = 0;
    st->diverged_index = -1;
    st->max_update = 0.0;
    st->energy = 0.0;
    st->active = 0;
    for (int z = 0; z < g->n[2]; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                double d = kfo_distance_idx_disp(live, u, g, x, y, z);
                if (d > p->max_surface - p->epsilon || d < -p->max_surface + p->epsilon)
                    continue;
                st->active++;
                int64_t i = IDX(g, x, y, z);
                int it = 0;
                double gr[3], gn;
                do {
                    kfo_energy_gradient(live, can, u, g, p, x, y, z, gr);
                    double du[3] = {-p->alpha * gr[0], -p->alpha * gr[1], -p->alpha * gr[2]};
                    u[i] += du[0];
                    u[n + i] += du[1];
                    u[2 * n + i] += du[2];
                    double dn = sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
                    if (dn > st->max_update)
                        st->max_update = dn;
                    if (dn <= p->gs_threshold)
                        break;
                    if (!isfinite(u[i]) || !isfinite(u[n + i]) || !isfinite(u[2 * n + i])) {
                        st->diverged_index = i;
                        return 1;
                    }
                    it += 1;
                    gn = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
                } while (gn > p->gs_threshold && it < p->max_iters);
                if (it > st->iters)
                    st->iters = it;
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a2: projective integration (SDF.cpp:137-197)                              */
/* ------------------------------------------------------------------------ */

/* General 4x4 inverse (row-major) by cofactors; returns 0 if singular.  The
 * reference calls Eigen's inverse() (SDF.cpp:147); for the rigid poses used
 * here both are exact to the last bit on identity. */
int kfo_invert4(const double m[16], double out[16])
{
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0)
        return 0;
    for (int i = 0; i < 16; i++)
        out[i] = inv[i] / det;
    return 1;
}

/* SDF::integrateDepthFrame, SDF.cpp:137-197.  depth is row-major H x W in
 * metres (the reference's CV_64F frame; float32 here, promoted to double).
 * Pixel pick: roundf((float)(p + 0.5)) -- the double is narrowed to float. */
void kfo_integrate(double *phi, int64_t *w, const kfo_grid *g, double clip,
                   const float *depth, int width, int height,
                   const double K[9], const double C2W[16], double min_depth, double max_depth)
{
    double W2C[16];
    if (!kfo_invert4(C2W, W2C))
        return;
    for (int z = 0; z < g->n[2]; z++) {
        double Z = g->voxel * (z + 0.5) + g->min[2];
        for (int y = 0; y < g->n[1]; y++) {
            double Y = g->voxel * (y + 0.5) + g->min[1];
            for (int x = 0; x < g->n[0]; x++) {
                double X = g->voxel * (x + 0.5) + g->min[0];
                double v[4] = {X, Y, Z, 1.0};
                double c[4];
                for (int r = 0; r < 4; r++)
                    c[r] = W2C[r * 4 + 0] * v[0] + W2C[r * 4 + 1] * v[1] + W2C[r * 4 + 2] * v[2] + W2C[r * 4 + 3] * v[3];
                double q[3] = {c[0] / c[2], c[1] / c[2], 1.0};
                double px = K[0] * q[0] + K[1] * q[1] + K[2] * q[2];
                double py = K[3] * q[0] + K[4] * q[1] + K[5] * q[2];
                int col = (int)roundf((float)(px + 0.5));
                if (col < 0 || col >= width)
                    continue;
                int row = (int)roundf((float)(py + 0.5));
                if (row < 0 || row >= height)
                    continue;
                double zc = c[2];
                double d = (double)depth[(int64_t)row * width + col];
                if (d < min_depth || d > max_depth)
                    continue;
                double dist = d - zc;
                if (dist < -clip)
                    continue;
                int64_t i = IDX(g, x, y, z);
                phi[i] = (phi[i] * (double)w[i] + dist) / (double)(w[i] + 1);
                w[i] += 1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a18-a20: warp and fuse                                                    */
/* ------------------------------------------------------------------------ */

/* SDF::update(disp), SDF.cpp:300-328: new phi/w at idx are the trilinear reads
 * at idx + 0.5 + u[idx]; the interpolated weight is truncated to long. */
void kfo_warp(double *phi, int64_t *w, const double *u, const kfo_grid *g)
{
    int64_t n = NVOX(g);
    double *np = (double *)malloc(sizeof(double) * (size_t)n);
    int64_t *nw = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->n[2]; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                double d[3];
                kfo_disp_at(u, g, x, y, z, d);
                double loc[3] = {(double)x + 0.5 + d[0], (double)y + 0.5 + d[1], (double)z + 0.5 + d[2]};
                double ww = kfo_get_weight(w, g, loc);
                double dd = kfo_get_distance(phi, g, loc);
                int64_t i = IDX(g, x, y, z);
                np[i] = dd;
                nw[i] = (int64_t)ww;
            }
    memcpy(phi, np, sizeof(double) * (size_t)n);
    memcpy(w, nw, sizeof(int64_t) * (size_t)n);
    free(np);
    free(nw);
}

/* SDF::fuse(other), SDF.cpp:199-241 with FUSE_BY_MERGE = true (config.cpp:7). */
void kfo_fuse_mode(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2, int64_t n, int merge)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t b = w2[i];
        if (b == 0)
            continue;
        int64_t a = w1[i];
        double d2 = phi2[i];
        if (a == 0) {
            phi1[i] = d2;
            w1[i] = b;
        } else if (merge) {
            double d1 = phi1[i];
            phi1[i] = ((double)a * d1 + (double)b * d2) / (double)(a + b);
            w1[i] = a + b;
        } else if (!(fabs(phi1[i]) < fabs(d2))) {
            /* FUSE_BY_MERGE false (SDF.cpp:229-236): the value nearer the surface wins */
            phi1[i] = d2;
            w1[i] = b;
        }
    }
}

void kfo_fuse(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2, int64_t n)
{
    kfo_fuse_mode(phi1, w1, phi2, w2, n, 1);
}

void kfo_fuse_warped_mode(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2,
                          const double *u, const kfo_grid *g, int merge)
{
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->n[2]; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                double d[3];
                kfo_disp_at(u, g, x, y, z, d);
                double loc[3] = {(double)x + 0.5 + d[0], (double)y + 0.5 + d[1], (double)z + 0.5 + d[2]};
                double b = kfo_get_weight(w2, g, loc);
                if (fabs(b) < (double)1e-5f)
                    continue;
                int64_t i = IDX(g, x, y, z);
                int64_t a = w1[i];
                double d2 = kfo_get_distance(phi2, g, loc);
                if (a == 0) {
                    phi1[i] = d2;
                    w1[i] = (int64_t)b;
                } else if (merge) {
                    double d1 = phi1[i];
                    phi1[i] = ((double)a * d1 + b * d2) / ((double)a + b);
                    w1[i] = (int64_t)((double)a + b);
                } else if (!(fabs(phi1[i]) < fabs(d2))) {
                    /* FUSE_BY_MERGE false (SDF.cpp:285-295) */
                    phi1[i] = d2;
                    w1[i] = (int64_t)b;
                }
            }
}

void kfo_fuse_warped(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2,
                     const double *u, const kfo_grid *g)
{
    kfo_fuse_warped_mode(phi1, w1, phi2, w2, u, g, 1);
}

/* ------------------------------------------------------------------------ */
/* a21-a23: marching cubes                                                   */
/* ------------------------------------------------------------------------ */

/* Edge e of Polygonise joins corners KFO_EDGE_ENDS[e] (the VertexInterp calls of
 * MarchingCubes.h:386-409: p1 is the first corner, p2 the second). */
static const int KFO_EDGE_ENDS[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                         {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

/* edgeTable[cube_index], MarchingCubes.h:26-59. */
int kfo_mc_edge_mask(int ci) { return kfo_edge_table[ci & 255]; }

/* triTable[cube_index] (MarchingCubes.h:61-318): 16 edge ids, -1 padded. */
void kfo_mc_case(int ci, int32_t out[16])
{
    for (int k = 0; k < 16; k++)
        out[k] = kfo_tri_table[ci & 255][k];
}

/* Polygonise + VertexInterp, MarchingCubes.h:326-422. */
int kfo_polygonise(const double p[8][3], const double val[8], double iso, double tris[5][3][3])
{
    int ci = 0;
    for (int k = 0; k < 8; k++)
        if (val[k] < iso)
            ci |= 1 << k;
    const int em = kfo_edge_table[ci];
    if (em == 0)
        return 0;
    double vl[12][3];
    for (int e = 0; e < 12; e++) {
        if (!(em & (1 << e)))
            continue;
        const int a = KFO_EDGE_ENDS[e][0], b = KFO_EDGE_ENDS[e][1];
        const double s = (iso - val[a]) / (val[a] - val[b]);
        for (int c = 0; c < 3; c++)
            vl[e][c] = p[a][c] + s * (p[a][c] - p[b][c]);
    }
    int nt = 0;
    for (int i = 0; kfo_tri_table[ci][i] != -1; i += 3) {
        for (int v = 0; v < 3; v++)
            for (int c = 0; c < 3; c++)
                tris[nt][v][c] = vl[kfo_tri_table[ci][i + v]][c];
        nt++;
    }
    return nt;
}

/* SDF::getMesh + ProcessVolumeCell, SDF.cpp:330-378, 396-414; vertices appended
 * with y negated (SimpleMesh.h:66-72), faces implicit (3t, 3t+1, 3t+2). */
int64_t kfo_marching_cubes(const double *phi, const kfo_grid *g, double max_surface, double clip,
                           double iso, double *verts, int64_t cap)
{
    static const int CORNER[8][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                     {1, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    double half[3] = {(double)g->n[0] / 2, (double)g->n[1] / 2, (double)g->n[2] / 2};
    int64_t count = 0;
    for (int z = 0; z < g->n[2]; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                double d = kfo_distance_at_index(phi, g, x, y, z);
                if (d > max_surface || d < -clip)
                    continue;
                double p[8][3], val[8];
                for (int k = 0; k < 8; k++) {
                    int cx = x + CORNER[k][0], cy = y + CORNER[k][1], cz = z + CORNER[k][2];
                    p[k][0] = (double)cx / half[0] - 1;
                    p[k][1] = (double)cy / half[1] - 1;
                    p[k][2] = (double)cz / half[2] - 1;
                    val[k] = kfo_distance_at_index(phi, g, cx, cy, cz);
                }
                double tris[5][3][3];
                int nt = kfo_polygonise(p, val, iso, tris);
                for (int t = 0; t < nt; t++) {
                    if (verts && count < cap) {
                        double *o = verts + count * 9;
                        for (int v = 0; v < 3; v++) {
                            o[v * 3 + 0] = tris[t][v][0];
                            o[v * 3 + 1] = -tris[t][v][1];
                            o[v * 3 + 2] = tris[t][v][2];
                        }
                    }
                    count++;
                }
            }
    return count;
}

/* ======================================================================== */
/* Adjoint formulation (no reference counterpart).                          */
/*                                                                          */
/* Notation: j integer voxel, q_j = j + u_j (index space), T = trilinear     */
/* interpolant of phi_live with the reference's cell-centred convention      */
/* (kfo_get_distance at q + 0.5: truncating base, OOB corners = Max+eps),    */
/* psi_j = T(q_j), g_j = grad T(q_j), H_j = Hessian of T at q_j (the          */
/* trilinear has zero diagonal second derivatives, so only the mixed ones).  */
/* Active mask m_j: -clip <= psi_j <= Max - eps (KillingFusion.cpp:273),     */
/* held fixed while differentiating.                                          */
/*                                                                          */
/* E(u) = sum_j m_j [ use_data * (psi_j - phi_can_j)^2 / (2 V)               */
/*                   + use_ls * omega_s * (s_j - 1)^2 / 2 ]                  */
/*      + use_killing * omega_k * sum_{i in [-1, n-1]^3}                      */
/*            ( ||D u_i||_F^2 + gamma * tr((D u_i)^2) )                      */
/* with s_j = sqrt(|g_j|^2 + eps^2), (D u_i)_{ab} = u^a_{i+e_b} - u^a_i and  */
/* u = 0 outside the grid.                                                   */
/*                                                                          */
/* dE/du_j = m_j [ (psi_j - phi_can_j)/V g_j + omega_s (s_j-1)/s_j H_j g_j ]  */
/*          - 2 omega_k [ sum_b (u_{j+e_b} - 2u_j + u_{j-e_b})               */
/*                      + gamma sum_b (D_a u^b_j - D_a u^b_{j-e_b}) ]_a      */
/* The data and level-set parts are the reference's formulas                 */
/* (KillingFusion.cpp:444-476) with the finite differences replaced by the   */
/* exact derivatives of the interpolant at the data-term sample; the Killing */
/* part is the exact gradient of the discrete Killing energy                 */
/* (DisplacementField.cpp:171-193 with forward differences), which equals     */
/* -2 (Lap u + gamma grad div u) as in DisplacementField.cpp:333-337 without */
/* the clamp and without the 1/(2h) factor the trilinear samples introduce.  */
/* ======================================================================== */

typedef struct adj_sample {
    double psi, g[3], hxy, hxz, hyz;
} adj_sample;

static void adj_sample_at(const double *live, const kfo_grid *g, double qx, double qy, double qz, adj_sample *s)
{
    int bx = (int)qx, by = (int)qy, bz = (int)qz;
    double fx = qx - (double)bx, fy = qy - (double)by, fz = qz - (double)bz;
    double v000 = kfo_distance_at_index(live, g, bx, by, bz);
    double v001 = kfo_distance_at_index(live, g, bx + 1, by, bz);
    double v010 = kfo_distance_at_index(live, g, bx, by + 1, bz);
    double v011 = kfo_distance_at_index(live, g, bx + 1, by + 1, bz);
    double v100 = kfo_distance_at_index(live, g, bx, by, bz + 1);
    double v101 = kfo_distance_at_index(live, g, bx + 1, by, bz + 1);
    double v110 = kfo_distance_at_index(live, g, bx, by + 1, bz + 1);
    double v111 = kfo_distance_at_index(live, g, bx + 1, by + 1, bz + 1);
    s->psi = kfo_interp3(v000, v001, v010, v011, v100, v101, v110, v111, fx, fy, fz);
    /* x-differences on the four x-edges (y0z0, y1z0, y0z1, y1z1) */
    double dx00 = v001 - v000, dx10 = v011 - v010, dx01 = v101 - v100, dx11 = v111 - v110;
    double dy00 = v010 - v000, dy10 = v011 - v001, dy01 = v110 - v100, dy11 = v111 - v101;
    double dz00 = v100 - v000, dz10 = v101 - v001, dz01 = v110 - v010, dz11 = v111 - v011;
    /* dy?? indexed (x, z); dz?? indexed (x, y) */
    s->g[0] = lerp1(lerp1(dx00, dx10, fy), lerp1(dx01, dx11, fy), fz);
    s->g[1] = lerp1(lerp1(dy00, dy10, fx), lerp1(dy01, dy11, fx), fz);
    s->g[2] = lerp1(lerp1(dz00, dz10, fx), lerp1(dz01, dz11, fx), fy);
    s->hxy = lerp1(dx10 - dx00, dx11 - dx01, fz);
    s->hxz = lerp1(dx01 - dx00, dx11 - dx10, fy);
    s->hyz = lerp1(dy01 - dy00, dy11 - dy10, fx);
}

static inline double uget(const double *u, const kfo_grid *g, int c, int x, int y, int z)
{
    if (!in_grid(g, x, y, z))
        return 0.0;
    return u[(int64_t)c * NVOX(g) + IDX(g, x, y, z)];
}

/* Local data + level-set terms at voxel j; returns the active flag. */
static int adj_local(const double *live, const double *can, const double *u, const kfo_grid *g,
                     const kfo_params *p, int x, int y, int z, double *e_local, double grad[3])
{
    int64_t n = NVOX(g), i = IDX(g, x, y, z);
    adj_sample s;
    adj_sample_at(live, g, (double)x + u[i], (double)y + u[n + i], (double)z + u[2 * n + i], &s);
    grad[0] = grad[1] = grad[2] = 0.0;
    *e_local = 0.0;
    if (!jacobi_active(s.psi, p))
        return 0;
    if (p->use_data) {
        double r = s.psi - can[i];
        *e_local += r * r / (2 * p->voxel_size);
        double c = r / p->voxel_size;
        for (int a = 0; a < 3; a++)
            grad[a] += c * s.g[a];
    }
    if (p->use_levelset) {
        double s2 = s.g[0] * s.g[0] + s.g[1] * s.g[1] + s.g[2] * s.g[2] + p->epsilon * p->epsilon;
        double sn = sqrt(s2);
        *e_local += p->omega_s * 0.5 * (sn - 1) * (sn - 1);
        double c = p->omega_s * (sn - 1) / sn;
        double hg0 = s.hxy * s.g[1] + s.hxz * s.g[2];
        double hg1 = s.hxy * s.g[0] + s.hyz * s.g[2];
        double hg2 = s.hxz * s.g[0] + s.hyz * s.g[1];
        grad[0] += c * hg0;
        grad[1] += c * hg1;
        grad[2] += c * hg2;
    }
    return 1;
}

/* Killing energy terms owned by voxel j: ||D u_j||^2 + gamma tr((D u_j)^2), plus
 * the terms of the i = j - e_c ghost voxels when j_c = 0. */
static double adj_killing_energy_at(const double *u, const kfo_grid *g, double gamma, int x, int y, int z)
{
    int j[3] = {x, y, z};
    double D[3][3]; /* D[a][b] = u^a_{j+e_b} - u^a_j */
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            int q[3] = {x, y, z};
            q[b] += 1;
            D[a][b] = uget(u, g, a, q[0], q[1], q[2]) - uget(u, g, a, x, y, z);
        }
    double e = 0.0, t = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            e += D[a][b] * D[a][b];
            t += D[a][b] * D[b][a];
        }
    e += gamma * t;
    for (int c = 0; c < 3; c++) {
        if (j[c] != 0)
            continue;
        double s = 0.0;
        for (int a = 0; a < 3; a++) {
            double v = uget(u, g, a, x, y, z);
            s += v * v;
        }
        double vc = uget(u, g, c, x, y, z);
        e += s + gamma * vc * vc;
    }
    return e;
}

static void adj_killing_grad_at(const double *u, const kfo_grid *g, double gamma, int x, int y, int z, double out[3])
{
    for (int a = 0; a < 3; a++) {
        double u0 = uget(u, g, a, x, y, z);
        double lap = 0.0;
        for (int b = 0; b < 3; b++) {
            int p[3] = {x, y, z}, m[3] = {x, y, z};
            p[b] += 1;
            m[b] -= 1;
            lap += uget(u, g, a, p[0], p[1], p[2]) - 2 * u0 + uget(u, g, a, m[0], m[1], m[2]);
        }
        double dv = 0.0;
        for (int b = 0; b < 3; b++) {
            /* D_a u^b at j and at j - e_b */
            int ja[3] = {x, y, z};
            ja[a] += 1;
            double d_j = uget(u, g, b, ja[0], ja[1], ja[2]) - uget(u, g, b, x, y, z);
            int jm[3] = {x, y, z};
            jm[b] -= 1;
            int jma[3] = {jm[0], jm[1], jm[2]};
            jma[a] += 1;
            double d_m = uget(u, g, b, jma[0], jma[1], jma[2]) - uget(u, g, b, jm[0], jm[1], jm[2]);
            dv += d_j - d_m;
        }
        out[a] = -2 * (lap + gamma * dv);
    }
}

double kfo_adj_energy(const double *live, const double *can, const double *u, const kfo_grid *g,
                      const kfo_params *p, double *e_voxel, int64_t *active)
{
    int nz = g->n[2];
    double *zs = (double *)calloc((size_t)nz, sizeof(double));
    int64_t *za = (int64_t *)calloc((size_t)nz, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic)
    for (int z = 0; z < nz; z++) {
        double s = 0.0;
        int64_t act = 0;
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                double el, gr[3];
                act += adj_local(live, can, u, g, p, x, y, z, &el, gr);
                double e = el;
                if (p->use_killing)
                    e += p->omega_k * adj_killing_energy_at(u, g, p->gamma, x, y, z);
                if (e_voxel)
                    e_voxel[IDX(g, x, y, z)] = e;
                s += e;
            }
        zs[z] = s;
        za[z] = act;
    }
    double tot = 0.0;
    int64_t act = 0;
    for (int z = 0; z < nz; z++) {
        tot += zs[z];
        act += za[z];
    }
    if (active)
        *active = act;
    free(zs);
    free(za);
    return tot;
}

double kfo_adj_energy_local(const double *live, const double *can, const double *u, const kfo_grid *g,
                            const kfo_params *p, int x, int y, int z)
{
    const int off[4][3] = {{0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    double e = 0.0;
    for (int k = 0; k < 4; k++) {
        int q[3] = {x + off[k][0], y + off[k][1], z + off[k][2]};
        if (q[0] < 0 || q[1] < 0 || q[2] < 0)
            continue;
        double el, gr[3];
        adj_local(live, can, u, g, p, q[0], q[1], q[2], &el, gr);
        e += el;
        if (p->use_killing)
            e += p->omega_k * adj_killing_energy_at(u, g, p->gamma, q[0], q[1], q[2]);
    }
    return e;
}

void kfo_adj_gradient(const double *live, const double *can, const double *u, const kfo_grid *g,
                      const kfo_params *p, double *grad, uint8_t *mask)
{
    int64_t n = NVOX(g);
#pragma omp parallel for schedule(dynamic)
    for (int z = 0; z < g->n[2]; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                int64_t i = IDX(g, x, y, z);
                double el, gr[3], kg[3] = {0, 0, 0};
                int act = adj_local(live, can, u, g, p, x, y, z, &el, gr);
                if (p->use_killing)
                    adj_killing_grad_at(u, g, p->gamma, x, y, z, kg);
                for (int a = 0; a < 3; a++)
                    grad[(int64_t)a * n + i] = gr[a] + p->omega_k * kg[a];
                if (mask)
                    mask[i] = (uint8_t)act;
            }
}

/* kfo_adj_gradient restricted to what a Jacobi step uses: the mask everywhere and the
 * gradient at active voxels (zero elsewhere).  A voxel with u_j = 0 samples the live SDF at
 * the node j itself (zero interpolation weights), so psi_j = live[j] exactly and an
 * out-of-band live[j] makes it inactive without the sample.  Returns the active count. */
static int64_t adj_gradient_active(const double *live, const double *can, const double *u, const kfo_grid *g,
                                   const kfo_params *p, double *grad, uint8_t *mask)
{
    int64_t n = NVOX(g);
    int nz = g->n[2];
    int64_t *za = (int64_t *)calloc((size_t)nz, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < g->n[1]; y++)
            for (int x = 0; x < g->n[0]; x++) {
                int64_t i = IDX(g, x, y, z);
                double el, gr[3] = {0, 0, 0}, kg[3] = {0, 0, 0};
                int act;
                if (u[i] == 0.0 && u[n + i] == 0.0 && u[2 * n + i] == 0.0 && !jacobi_active(live[i], p))
                    act = 0;
                else
                    act = adj_local(live, can, u, g, p, x, y, z, &el, gr);
                if (act && p->use_killing)
                    adj_killing_grad_at(u, g, p->gamma, x, y, z, kg);
                for (int a = 0; a < 3; a++)
                    grad[(int64_t)a * n + i] = act ? gr[a] + p->omega_k * kg[a] : 0.0;
                mask[i] = (uint8_t)act;
                za[z] += act;
            }
    int64_t act = 0;
    for (int z = 0; z < nz; z++)
        act += za[z];
    free(za);
    return act;
}

/* Jacobi gradient descent on E: active voxels move by -alpha dE/du. */
int kfo_adj_deform(const double *live, const double *can, double *u, const kfo_grid *g,
                   const kfo_params *p, kfo_stats *st)
{
    int64_t n = NVOX(g);
    double *grad = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    uint8_t *mask = (uint8_t *)malloc((size_t)n);
    st->iters = 0;
    st->diverged_index = -1;
    st->max_update = 0.0;
    st->energy = 0.0;
    st->active = 0;
    for (int it = 0; it < p->max_iters; it++) {
        const int64_t act = adj_gradient_active(live, can, u, g, p, grad, mask);
        double m = 0.0;
        for (int64_t i = 0; i < n; i++) {
            if (!mask[i])
                continue;
            double du[3];
            for (int a = 0; a < 3; a++) {
                du[a] = -p->alpha * grad[(int64_t)a * n + i];
                u[(int64_t)a * n + i] = u[(int64_t)a * n + i] + du[a];
            }
            double nn = sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
            if (nn > m)
                m = nn;
            if (st->diverged_index < 0 && (!isfinite(u[i]) || !isfinite(u[n + i]) || !isfinite(u[2 * n + i])))
                st->diverged_index = i;
        }
        st->iters = it + 1;
        st->max_update = m;
        st->active = act;
        if (st->diverged_index >= 0)
            break;
        if (p->stop_max_update > 0.0 && m < p->stop_max_update)
            break;
    }
    if (st->diverged_index < 0)
        st->energy = kfo_adj_energy(live, can, u, g, p, NULL, NULL);
    free(grad);
    free(mask);
    return st->diverged_index >= 0;
}

/* Threads the OpenMP loops of this library use (for reporting the CPU baseline). */
#ifdef _OPENMP
#include <omp.h>
#endif
int kfo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void kfo_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}
