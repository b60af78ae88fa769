/*
 * kf_oracle.h -- CPU restatement (FP64, plain C99 + OpenMP) of KillingFusion's
 * per-frame volumetric hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP library
 * (killingfusion_amd/libkf.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only to compare against or to time
 * as the CPU baseline.  The product path never calls into it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root, all under KillingFusion/).  Reference
 * semantics are reproduced literally, including the quirks:
 *   - truncating int casts for trilinear base indices (SDF.cpp:509,590,
 *     DisplacementField.cpp:43), so locations in (-1,0) extrapolate;
 *   - the -0.5 cell-centre offset for phi reads only (SDF.cpp:506,587);
 *   - the linear-index-only bounds test for displacement reads, so x=nx or y=ny
 *     wraps into the next row/slice (DisplacementField.cpp:24-26,32-34);
 *   - the extra +0.5 in getDistancef that puts gradient/Hessian samples half a
 *     voxel off the data-term sample (SDF.cpp:577-582,726-744,842-886);
 *   - roundf on a float-narrowed pixel coordinate in integration (SDF.cpp:167,171);
 *   - the Killing-gradient clamp applied before the 1/(4h^2) division
 *     (DisplacementField.cpp:339-346);
 *   - double->long truncation of interpolated weights in warp/fuse
 *     (SDF.cpp:316-320, 258-278).
 *
 * Data layout (same as the device library): x fastest, index z*nx*ny + y*nx + x
 * (SDF.cpp:27,189; DisplacementField.cpp:15,32).  Displacements are SoA:
 * u[0..N) = ux, u[N..2N) = uy, u[2N..3N) = uz, in voxel units.  Weights are int64
 * (the reference's `long`, SDF.h:25).
 *
 * Pins (tests/test_oracle_kat.py): every function is checked against the reference's own
 * self-test expectations, closed forms, brute force or an independent transcription, as
 * listed per function there.  Parity unpinned: results that depend on Eigen's evaluation
 * order -- the K (X/Z, Y/Z, 1) projection and the 4x4 inverse of integration (SDF.cpp:146-164),
 * norm() in the Killing clamp (DisplacementField.cpp:343-344) and the level-set term
 * (KillingFusion.cpp:474) -- are restated in plain left-to-right FP64 and can differ from the
 * reference binary in the last bits; nothing at the stated tolerances depends on them.
 *
 * The second half of this file is the FP64 oracle of the adjoint formulation
 * (a discrete energy and its exact gradient) that the device library's fused
 * forward/backward kernels implement.  It has no counterpart in the reference;
 * its definition is stated next to the functions.
 */
#ifndef KF_ORACLE_H
#define KF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kfo_grid {
    int32_t n[3];      /* dims x, y, z */
    double min[3];     /* world position of the grid's min corner (m) */
    double voxel;      /* voxel edge (m) */
    double oob;        /* phi returned outside the grid: MaxSurfaceVoxelDistance + epsilon */
} kfo_grid;

typedef struct kfo_params {
    double voxel_size;     /* global VoxelSize used by the data term (config.cpp:4) */
    double max_surface;    /* MaxSurfaceVoxelDistance (config.cpp:6) */
    double clip;           /* UnknownClipDistance (config.cpp:5) */
    double epsilon;        /* epsilon (config.cpp:57) */
    double alpha;          /* gradient step (config.cpp:45) */
    double omega_k;        /* omegaKilling (config.cpp:48) */
    double gamma;          /* gammaKilling (config.cpp:51) */
    double omega_s;        /* omegaLevelSet (config.cpp:54) */
    double h;              /* deltaSize, voxel units (config.cpp:29) */
    double killing_clamp;  /* 0.1, hard-coded at DisplacementField.cpp:342 */
    int32_t use_data, use_levelset, use_killing;   /* EnergyTypeUsed (config.cpp:23) */
    int32_t max_iters;                             /* KILLING_MAX_ITERATIONS (config.cpp:42) */
    double stop_max_update;  /* stop once max |du| < this (voxel); <= 0 disables */
    double gs_threshold;     /* `threshold` of the per-voxel loop (config.cpp:43) */
    int32_t use_trust;       /* UseTrustStrategy (config.cpp:27), honoured with the data energy alone */
    int32_t pad;
} kfo_params;

typedef struct kfo_stats {
    int32_t iters;
    int64_t diverged_index;  /* -1 if all finite */
    double max_update;       /* max |du| of the last iteration (voxel) */
    double energy;           /* adjoint only: energy of the final field */
    int64_t active;          /* active voxels in the last iteration */
} kfo_stats;

/* ---- a1: grid sizing (SDF.cpp:117-135) ---- */
void kfo_dims_from_bounds(const double min[3], const double max[3], double voxel, int32_t out[3]);
void kfo_sdf_init(double *phi, int64_t *w, int64_t n, double oob);

/* ---- a3-a7: reads ---- */
double kfo_interp3(double v000, double v001, double v010, double v011,
                   double v100, double v101, double v110, double v111,
                   double x, double y, double z);
double kfo_distance_at_index(const double *phi, const kfo_grid *g, int x, int y, int z);
int64_t kfo_weight_at_index(const int64_t *w, const kfo_grid *g, int x, int y, int z);
double kfo_get_distance(const double *phi, const kfo_grid *g, const double loc[3]);
double kfo_get_weight(const int64_t *w, const kfo_grid *g, const double loc[3]);
void kfo_disp_at(const double *u, const kfo_grid *g, int x, int y, int z, double out[3]);
void kfo_disp_at_f(const double *u, const kfo_grid *g, const double loc[3], double out[3]);
double kfo_distance_idx_disp(const double *phi, const double *u, const kfo_grid *g, int x, int y, int z);
double kfo_distance_f(const double *phi, const double *u, const kfo_grid *g, const double loc[3]);

/* ---- a8, a9, a12: stencils ---- */
void kfo_distance_gradient_disp(const double *phi, const double *u, const kfo_grid *g,
                                int x, int y, int z, double h, double out[3]);
void kfo_distance_hessian_disp(const double *phi, const double *u, const kfo_grid *g,
                               int x, int y, int z, double h, double out[9]);
int kfo_distance_gradient_plain(const double *phi, const kfo_grid *g, const double loc[3], double out[3]);
int kfo_distance_hessian_plain(const double *phi, const kfo_grid *g, const double loc[3], double h, double out[9]);

/* ---- a10, a11: displacement-field stencils ---- */
void kfo_jacobian(const double *u, const kfo_grid *g, double x, double y, double z, double h, double out[9]);
double kfo_killing_energy(const double *u, const kfo_grid *g, double x, double y, double z,
                          double h, double gamma);
void kfo_killing_gradient2(const double *u, const kfo_grid *g, int x, int y, int z,
                           double h, double gamma, double clamp, double out[3]);

/* ---- a13-a16: energy gradient ---- */
void kfo_data_gradient(const double *live, const double *can, const double *u, const kfo_grid *g,
                       const kfo_params *p, int x, int y, int z, double out[3]);
void kfo_levelset_gradient(const double *live, const double *u, const kfo_grid *g,
                           const kfo_params *p, int x, int y, int z, double out[3]);
void kfo_energy_gradient(const double *live, const double *can, const double *u, const kfo_grid *g,
                         const kfo_params *p, int x, int y, int z, double out[3]);
/* Whole-grid gradient in SoA (zero where the voxel is skipped when masked != 0). */
void kfo_energy_gradient_grid(const double *live, const double *can, const double *u,
                              const kfo_grid *g, const kfo_params *p, int masked, double *grad);

/* ---- a17: optimiser ---- */
int kfo_deform_jacobi(const double *live, const double *can, double *u, const kfo_grid *g,
                      const kfo_params *p, kfo_stats *st);
int kfo_deform_gs_serial(const double *live, const double *can, double *u, const kfo_grid *g,
                         const kfo_params *p, kfo_stats *st);

/* ---- a2: integration ---- */
void kfo_integrate(double *phi, int64_t *w, const kfo_grid *g, double clip,
                   const float *depth, int width, int height,
                   const double K[9], const double C2W[16], double min_depth, double max_depth);
int kfo_invert4(const double m[16], double out[16]);

/* ---- a18-a20: warp and fuse ---- */
void kfo_warp(double *phi, int64_t *w, const double *u, const kfo_grid *g);
void kfo_fuse(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2, int64_t n);
/* SDF::fuse with FUSE_BY_MERGE = merge (config.cpp:7; SDF.cpp:199-241): 0 keeps, per voxel, the
 * value nearer the surface with its own weight. */
void kfo_fuse_mode(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2, int64_t n, int merge);
void kfo_fuse_warped(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2,
                     const double *u, const kfo_grid *g);
/* SDF::fuse(other, disp) with FUSE_BY_MERGE = merge (SDF.cpp:244-298). */
void kfo_fuse_warped_mode(double *phi1, int64_t *w1, const double *phi2, const int64_t *w2,
                          const double *u, const kfo_grid *g, int merge);

/* ---- a21-a23: marching cubes ---- */
/* Returns the triangle count; writes min(count, cap) triangles (9 doubles each,
 * y already negated as SimpleMesh::AddVertex does) when verts != NULL. */
int64_t kfo_marching_cubes(const double *phi, const kfo_grid *g, double max_surface, double clip,
                           double iso, double *verts, int64_t cap);
int kfo_polygonise(const double p[8][3], const double val[8], double iso, double tris[5][3][3]);
int kfo_mc_edge_mask(int cube_index);
void kfo_mc_case(int cube_index, int32_t out[16]);

/* ---- adjoint formulation (no reference counterpart) ---- */
double kfo_adj_energy(const double *live, const double *can, const double *u, const kfo_grid *g,
                      const kfo_params *p, double *e_voxel, int64_t *active);
/* The terms of kfo_adj_energy that depend on u at voxel (x, y, z): the per-voxel energies of
 * (x, y, z) and of its lower neighbours (x-1, y, z), (x, y-1, z), (x, y, z-1) (forward
 * differences reach (x, y, z) only from those).  Differences of it in u_(x,y,z) equal
 * differences of the total energy. */
double kfo_adj_energy_local(const double *live, const double *can, const double *u, const kfo_grid *g,
                            const kfo_params *p, int x, int y, int z);
void kfo_adj_gradient(const double *live, const double *can, const double *u, const kfo_grid *g,
                      const kfo_params *p, double *grad, uint8_t *mask);
int kfo_adj_deform(const double *live, const double *can, double *u, const kfo_grid *g,
                   const kfo_params *p, kfo_stats *st);

int kfo_num_threads(void);
/* Threads of the following parallel regions (the timing of the CPU baseline on 1 core). */
void kfo_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
