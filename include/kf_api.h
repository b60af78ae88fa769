/*
 * kf_api.h -- C ABI of the MI355X (gfx950) KillingFusion hot-path library
 * (killingfusion_amd/libkf.so).
 *
 * The reference has no FFI: the path sits behind the C++ classes SDF
 * (KillingFusion/include/SDF.h:16-200), DisplacementField
 * (KillingFusion/include/DisplacementField.h:9-66) and the private optimiser
 * methods of KillingFusion (KillingFusion/include/KillingFusion.h:28-58),
 * driven by KillingFusion::processNextFrame (KillingFusion/src/KillingFusion.cpp:109-175).
 * Each entry point below names the reference member it replaces.
 *
 * Conventions
 *  - Every call returns kf_status; no C++ exception crosses this boundary.
 *    kf_last_error() returns a message for the calling thread's last failure.
 *  - Handles are bound to their context's HIP stream.  Calls are stream-ordered;
 *    those that return host values (downloads, stats, mesh counts) synchronise
 *    the stream first.  Handles are not thread-safe.
 *  - Grids are dense, x fastest: index = z*nx*ny + y*nx + x
 *    (SDF.cpp:27,189; DisplacementField.cpp:15,32).
 *  - Signed distances are metres; displacements are voxel units
 *    (DisplacementField.h:12); weights are int32 (the reference uses long,
 *    SDF.h:25).
 *  - Host arrays are plain pointers; displacements cross the ABI as SoA float32
 *    [3][N] (ux plane, uy plane, uz plane).
 */
#ifndef KF_API_H
#define KF_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KF_API_VERSION 1

typedef enum kf_status {
    KF_OK = 0,
    KF_E_INVALID_ARG = 1,  /* null/short buffer, bad dims, mismatched shapes */
    KF_E_DIVERGED = 2,     /* non-finite displacement (KillingFusion.cpp:327-332) */
    KF_E_OOM = 3,          /* device allocation failed */
    KF_E_HIP = 4,          /* HIP runtime error */
    KF_E_NO_DEVICE = 5,    /* no usable gfx950 device */
    KF_E_CAPACITY = 6      /* output capacity too small (count still returned) */
} kf_status;

/* Storage type of a volume's signed distances.  F32 stores metres; BF16/F16
 * store phi / (max_surface + epsilon), so the untouched value Max+eps
 * (SDF.cpp:133) is exactly 1.0 in every format.  Arithmetic is FP32 in all. */
typedef enum kf_dtype { KF_F32 = 0, KF_BF16 = 1, KF_F16 = 2 } kf_dtype;

/* Energy-gradient formulation used by kf_deform. */
typedef enum kf_method {
    KF_METHOD_REFERENCE = 0, /* KillingFusion::computeEnergyGradient, FD stencils (KillingFusion.cpp:399-476) */
    KF_METHOD_ADJOINT = 1    /* exact gradient of the discrete energy of kf_energy_forward */
} kf_method;

/* Which voxels a Jacobi step visits, with either method (the field it produces is the same bit
 * for bit).
 * A step moves voxel j only when -clip <= phi_live(j + u_j) <= Max - eps
 * (KillingFusion.cpp:272-274), and with u_j = 0 that reads phi_live[j]; so the voxels that
 * can move during one kf_deform call are those with u_j != 0 or phi_live[j] in that band,
 * fixed at the start of the call.  The active-set sweep lists them once and steps only
 * them; the dense sweep streams the whole grid every iteration.  The active-set sweep runs on
 * grids of fewer than 2^32 voxels and lists at most 2^30 of them; the dense sweep runs otherwise,
 * whatever the sweep asked for. */
typedef enum kf_sweep {
    KF_SWEEP_AUTO = 0,   /* active set when it holds at most 1/8 of the voxels, else dense */
    KF_SWEEP_DENSE = 1,  /* every voxel every iteration */
    KF_SWEEP_ACTIVE = 2  /* the listed voxels only */
} kf_sweep;

typedef struct kf_context kf_context;
typedef struct kf_volume kf_volume;
typedef struct kf_field kf_field;
typedef struct kf_pipeline kf_pipeline;

/* Volume geometry: SDF::SDF(voxelSize, min, max, unknownClipDistance), SDF.cpp:14-30,
 * with the dims given explicitly (kf_dims_from_bounds reproduces SDF.cpp:122). */
typedef struct kf_volume_desc {
    int32_t dims[3];
    double min_corner[3];   /* m_min3dLoc */
    double voxel_size;      /* m_voxelSize */
    double clip_distance;   /* UnknownClipDistance (config.cpp:5) */
    double max_surface;     /* MaxSurfaceVoxelDistance (config.cpp:6) */
    double epsilon;         /* epsilon (config.cpp:57); untouched phi = max_surface + epsilon */
    int32_t dtype;          /* kf_dtype */
    int32_t reserved;
} kf_volume_desc;

/* Optimiser parameters: the compile-time globals of config.cpp:4-57 as a runtime struct. */
typedef struct kf_params {
    double voxel_size;      /* VoxelSize, divides the data residual (KillingFusion.cpp:454) */
    double max_surface;     /* MaxSurfaceVoxelDistance, skip test (KillingFusion.cpp:273) */
    double clip_distance;   /* UnknownClipDistance, skip test (KillingFusion.cpp:273) */
    double epsilon;         /* epsilon (config.cpp:57) */
    double alpha;           /* gradient step (config.cpp:45) */
    double omega_killing;   /* omegaKilling (config.cpp:48) */
    double gamma_killing;   /* gammaKilling (config.cpp:51) */
    double omega_levelset;  /* omegaLevelSet (config.cpp:54) */
    double delta_size;      /* deltaSize h, voxels (config.cpp:29) */
    double killing_clamp;   /* 0.1 (DisplacementField.cpp:342) */
    int32_t use_data, use_levelset, use_killing;  /* EnergyTypeUsed (config.cpp:23) */
    int32_t max_iters;      /* KILLING_MAX_ITERATIONS (config.cpp:42) */
    double stop_max_update; /* stop once max |du| < this (voxels); <= 0 runs max_iters */
    int32_t method;         /* kf_method */
    int32_t sweep;          /* kf_sweep (either method) */
    int32_t use_trust;      /* UseTrustStrategy (config.cpp:27, off): the reference method's trust region,
                               honoured with the data energy alone (KillingFusion.cpp:288-314) */
    int32_t reserved;       /* 0 */
} kf_params;

typedef struct kf_deform_stats {
    int32_t iters;           /* iterations run */
    int32_t sweep;           /* the kf_sweep that ran (KF_SWEEP_DENSE or KF_SWEEP_ACTIVE) */
    int64_t diverged_index;  /* first voxel with a non-finite displacement, or -1 */
    double max_update;       /* max |du| of the last iteration, voxels (FP64) */
    double energy;           /* adjoint: energy of the final field (FP64); reference: 0 */
    uint64_t active_voxels;  /* voxels updated in the last iteration */
    uint64_t listed_voxels;  /* voxels that could move (0 when they were not counted: a dense sweep) */
} kf_deform_stats;

/* ---- library / context ---- */
const char *kf_status_string(kf_status s);
const char *kf_last_error(void);
int kf_api_version(void);
/* Fills p with the reference defaults of config.cpp:4-57 (voxel 12 mm, Max = clip = 4 voxels,
 * alpha 0.025, omega_k 0.04, gamma 0, omega_s 0.1, h 0.05, 1024 iterations). */
kf_status kf_params_default(kf_params *p);
/* stream: a hipStream_t, or NULL to create one owned by the context. */
kf_status kf_context_create(int device, void *stream, kf_context **out);
kf_status kf_context_destroy(kf_context *ctx);
kf_status kf_context_synchronize(kf_context *ctx);
/* The hipStream_t the context's work is ordered on (for callers that time or chain work). */
void *kf_context_stream(kf_context *ctx);
/* Iteration loops of the active-set sweep (kf_deform and kf_batch_deform without a stop rule)
 * as captured hipGraphs (no reference counterpart): with enable != 0, a loop whose launch shape
 * recurs on consecutive calls is captured the second time and replayed after that; with 0 (the
 * default) every loop is launched directly.  Results are the same bit for bit either way; on
 * MI355X with ROCm 7.2 direct launches keep the GPU as busy as a replay and came out 1-3%
 * faster, so capture is for callers whose host cannot keep ahead of the GPU. */
kf_status kf_context_set_loop_capture(kf_context *ctx, int32_t enable);
/* Page-locked host memory (no reference counterpart): depth frames held in it reach the GPU
 * by DMA, without the staging copy a pageable buffer costs (kf_integrate_depth,
 * kf_pipeline_process_frame).  Free with kf_host_free. */
kf_status kf_host_alloc(size_t bytes, void **out);
kf_status kf_host_free(void *p);

/* ---- SDF volume (SDF.h:16-200) ---- */
/* SDF::computeVoxelGridSize, SDF.cpp:117-124: int((max-min)/voxel)+1 per axis. */
kf_status kf_dims_from_bounds(const double min_corner[3], const double max_corner[3], double voxel,
                              int32_t dims_out[3]);
/* SDF::SDF + allocateMemoryForSDF, SDF.cpp:14-30,126-135: phi = max_surface+epsilon, w = 0. */
kf_status kf_volume_create(kf_context *ctx, const kf_volume_desc *desc, kf_volume **out);
kf_status kf_volume_destroy(kf_volume *vol);
kf_status kf_volume_describe(const kf_volume *vol, kf_volume_desc *out);
/* SDF::allocateMemoryForSDF fill, SDF.cpp:133-134. */
kf_status kf_volume_reset(kf_volume *vol);
/* Sets m_voxelGridTSDF / m_voxelGridWeight from host arrays (phi metres; weight may be NULL = 0). */
kf_status kf_volume_upload(kf_volume *vol, const float *phi, const int32_t *weight);
/* Reads them back (either pointer may be NULL). */
kf_status kf_volume_download(const kf_volume *vol, float *phi, int32_t *weight);
/* SDF(const SDF&) + copy of the voxel data (same dims and dtype). */
kf_status kf_volume_copy(kf_volume *dst, const kf_volume *src);

/* ---- displacement field (DisplacementField.h:9-66) ---- */
/* DisplacementField::DisplacementField, DisplacementField.cpp:8-16 (zero field). */
kf_status kf_field_create(kf_context *ctx, const int32_t dims[3], kf_field **out);
kf_status kf_field_destroy(kf_field *f);
/* DisplacementField::initializeAllVoxels, DisplacementField.cpp:82-88. */
kf_status kf_field_fill(kf_field *f, const double value[3]);
kf_status kf_field_upload(kf_field *f, const float *soa);
kf_status kf_field_download(const kf_field *f, float *soa);
/* DisplacementField::operator+, DisplacementField.cpp:73-80 (dst += src). */
kf_status kf_field_add(kf_field *dst, const kf_field *src);

/* ---- per-frame kernels ---- */
/* SDF::integrateDepthFrame, SDF.cpp:137-197.  depth: host float32 [height][width], metres
 * (0 = invalid); K row-major 3x3; C2W row-major 4x4 camera-to-world pose. */
kf_status kf_integrate_depth(kf_volume *vol, const float *depth, int32_t width, int32_t height,
                             const double K[9], const double C2W[16], double min_depth,
                             double max_depth);
/* KillingFusion::computeEnergyGradient (KillingFusion.cpp:399-442) at every voxel, written to
 * grad_out (zero at voxels the optimiser would skip, KillingFusion.cpp:272-274). */
kf_status kf_energy_gradient_ref(const kf_volume *live, const kf_volume *can, const kf_field *u,
                                 const kf_params *p, kf_field *grad_out);
/* Forward pass of the adjoint formulation: total energy (FP64, fixed-order sum) and,
 * optionally, per-voxel energy (host float32 [N]).  No reference counterpart: the
 * reference evaluates the Killing energy density only in self-tests
 * (DisplacementField.cpp:171-193). */
kf_status kf_energy_forward(const kf_volume *live, const kf_volume *can, const kf_field *u,
                            const kf_params *p, double *energy_out, float *energy_per_voxel);
/* Backward pass: exact dE/du at every voxel (zero data/level-set part at inactive voxels). */
kf_status kf_energy_backward(const kf_volume *live, const kf_volume *can, const kf_field *u,
                             const kf_params *p, kf_field *grad_out);
/* KillingFusion::computeDisplacementField, all-voxels (Jacobi) mode
 * (KillingFusion.cpp:248-351): live = src, can = dest, u = srcToDest, updated in place.
 * p->method selects the reference gradient or the fused adjoint kernel; p->sweep selects the
 * dense or the active-set sweep (same field; see kf_sweep). */
kf_status kf_deform(const kf_volume *live, const kf_volume *can, kf_field *u, const kf_params *p,
                    kf_deform_stats *stats);
/* SDF::update(displacementField), SDF.cpp:300-328. */
kf_status kf_warp(kf_volume *vol, const kf_field *u);
/* The FUSE_BY_MERGE switch of the reference's fuse (config.cpp:7). */
typedef enum kf_fuse_mode_t {
    KF_FUSE_MERGE = 0,   /* weighted average, weights summed (FUSE_BY_MERGE true, the reference's value) */
    KF_FUSE_NEAREST = 1  /* the value nearer the surface wins, with its own weight (FUSE_BY_MERGE false) */
} kf_fuse_mode_t;
/* SDF::fuse(otherSdf), SDF.cpp:199-241 (FUSE_BY_MERGE). */
kf_status kf_fuse(kf_volume *can, const kf_volume *live);
/* The same with either branch of FUSE_BY_MERGE (a kf_fuse_mode_t). */
kf_status kf_fuse_mode(kf_volume *can, const kf_volume *live, int32_t mode);
/* SDF::fuse(otherSdf, displacementField), SDF.cpp:244-298. */
kf_status kf_fuse_warped(kf_volume *can, const kf_volume *live, const kf_field *u);
kf_status kf_fuse_warped_mode(kf_volume *can, const kf_volume *live, const kf_field *u, int32_t mode);
/* SDF::getMesh (SDF.cpp:396-414) + ProcessVolumeCell (SDF.cpp:330-378): triangle soup in the
 * reference's order, 9 floats per triangle (3 vertices, y negated as SimpleMesh::AddVertex,
 * SimpleMesh.h:66-72); faces are implicit (3t, 3t+1, 3t+2).  *n_tris always receives the
 * count; verts may be NULL (size query); KF_E_CAPACITY if cap_tris < count. */
kf_status kf_marching_cubes(const kf_volume *vol, double iso, uint64_t *n_tris, float *verts,
                            uint64_t cap_tris);
/* Welded (indexed) variant of kf_marching_cubes; no reference counterpart: the reference keeps
 * the soup, and its viewer computes one normal per face (mainGui.cpp:68-89).  One vertex per cut
 * grid edge that a visited cell uses, numbered in grid-edge order (3 float, y negated as the
 * soup); triangles in the soup's order as 3 vertex indices; normals (may be NULL) are the unit
 * SDF gradient at each vertex.  *n_verts / *n_tris always receive the counts; verts, normals and
 * tris may all be NULL (size query); KF_E_CAPACITY if a capacity is short. */
kf_status kf_marching_cubes_indexed(const kf_volume *vol, double iso, uint64_t *n_verts, uint64_t *n_tris,
                                    float *verts, float *normals, uint32_t *tris, uint64_t cap_verts,
                                    uint64_t cap_tris);

/* Largest grid (voxels) kf_marching_cubes_noalloc takes: its scratch holds the unit and block
 * totals, offsets and list of blocks with triangles of 2^27 voxels (2 MiB per GPU). */
#define KF_MC_NOALLOC_MAX_VOXELS (1LL << 27)
/* kf_marching_cubes into caller device memory, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid allocation).  Enqueues on the volume's context stream and returns
 * without waiting: *count (device memory, one uint64) receives the triangle count, verts (device
 * memory) the first min(count, cap_tris) triangles in kf_marching_cubes's order and layout;
 * cap_tris = 0 (verts may then be NULL) only counts.  Grids above KF_MC_NOALLOC_MAX_VOXELS are
 * rejected (KF_E_INVALID_ARG).  The totals, offsets and list live in one device array per GPU,
 * held until the process ends, that every call on that GPU shares: calls that overlap on one GPU
 * (other streams or threads) race and must be ordered by the caller. */
kf_status kf_marching_cubes_noalloc(const kf_volume *vol, double iso, float *verts, uint64_t cap_tris,
                                    uint64_t *count);

/* ---- rigid pre-registration (no reference counterpart: the reference leaves SDF-2-SDF
 * registration as a future task, KillingFusion.cpp:76,133) ---- */
typedef struct kf_rigid_stats {
    int32_t iters;           /* Gauss-Newton updates applied */
    int32_t reserved;
    uint64_t voxels;         /* canonical band voxels whose sample fell inside the live grid */
    double initial_energy;   /* 1/2 sum (phi_live(T x) - phi_can(x))^2 at the input pose (m^2) */
    double energy;           /* the same at the returned pose */
    double last_step;        /* |rotation| (rad) + |translation| / voxel of the last update */
} kf_rigid_stats;
/* Rigid pose T (row-major 4x4, canonical world -> live world) minimising the squared SDF
 * difference over the canonical truncation band, by Gauss-Newton from the given pose (in/out).
 * FP64 accumulation, reproducible. */
kf_status kf_rigid_align(const kf_volume *live, const kf_volume *can, int32_t max_iters, double pose[16],
                         kf_rigid_stats *stats);
/* u_j = T(x_j) - j in the live grid's index units for every voxel j of the canonical grid: the
 * displacement field of a rigid pose, a starting point for kf_deform. */
kf_status kf_field_from_rigid(kf_field *u, const kf_volume *can, const kf_volume *live, const double pose[16]);

/* ---- batched deform (independent volumes of uneven size) ----
 * No reference counterpart: the reference deforms one live SDF per call
 * (KillingFusion::computeDisplacementField, KillingFusion.cpp:240-397); this runs
 * kf_deform on count independent (live[b], can[b], u[b]) triplets with one set of params.
 * With the adjoint method and the active-set sweep (p->sweep), every iteration is one launch
 * over the lists of all volumes, each volume with its own stop rule.  With the dense sweep,
 * every iteration is one persistent launch over the tiles of all volumes, the iteration loop
 * is a captured hipGraph (replayed while the batch keeps its shape), and the tile shape is
 * the autotuner's choice for the batch's shape.  Results equal per-volume kf_deform bit for
 * bit.  stats (may be NULL) receives one entry per volume; energies are evaluated only when
 * stats is given.  The reference method runs volume by volume.  Returns KF_E_DIVERGED if any
 * volume diverged (the others still complete). */
kf_status kf_batch_deform(const kf_volume *const *live, const kf_volume *const *can,
                          kf_field *const *u, int32_t count, const kf_params *p,
                          kf_deform_stats *stats);

/* Tile shape of the batched launch. */
typedef struct kf_batch_config {
    int32_t voxels_per_thread;  /* a tile is 256 x this consecutive voxels of one volume (1, 2, 4 or 8) */
    int32_t blocks;             /* grid of the persistent launch (== tiles: one block per tile) */
    int32_t candidates;         /* shapes timed by the tuner */
    int32_t reserved;
    double us_per_iteration;    /* the chosen shape's time per launch on this batch */
} kf_batch_config;

/* Autotuner of kf_batch_deform (no reference counterpart): times every candidate shape on
 * this batch (launches that read u[b] and write only its scratch buffer, so the fields keep
 * their values) and caches the fastest for batches of the same dims and dtype.  force = 1
 * re-times a cached shape; force = 2 pins the shape given in *best (voxels_per_thread, blocks;
 * blocks 0 = one block per tile) without timing.  best may be NULL unless force = 2. */
kf_status kf_batch_autotune(const kf_volume *const *live, const kf_volume *const *can,
                            kf_field *const *u, int32_t count, const kf_params *p, int32_t force,
                            kf_batch_config *best);

/* ---- per-frame pipeline: KillingFusion::processNextFrame, KillingFusion.cpp:109-175 ---- */
typedef struct kf_pipeline_desc {
    kf_volume_desc volume;   /* canonical and live volume geometry */
    kf_params params;
    double K[9];
    double C2W[16];
    double min_depth, max_depth;
    int32_t width, height;
    int32_t fuse_mode;       /* kf_fuse_mode_t of the per-frame fuse (0: KF_FUSE_MERGE, the reference's) */
    int32_t reserved;
} kf_pipeline_desc;

typedef struct kf_frame_stats {
    int32_t frame_index;
    int32_t reserved;
    kf_deform_stats deform;
    double ms_integrate, ms_deform, ms_fuse, ms_total;  /* the reference's printf columns */
} kf_frame_stats;

/* The meshes KillingFusion::processNextFrame returns (KillingFusion.cpp:171-173), as bits. */
typedef enum kf_frame_mesh {
    KF_MESH_LIVE = 1,         /* the frame's live SDF before the warp: currSdf->getMesh() (:157) */
    KF_MESH_LIVE_WARPED = 2,  /* the live SDF warped into the canonical frame (:160) */
    KF_MESH_CANONICAL = 4     /* the canonical SDF after the fuse: m_canonicalSdf->getMesh() (:170) */
} kf_frame_mesh;

kf_status kf_pipeline_create(kf_context *ctx, const kf_pipeline_desc *desc, kf_pipeline **out);
kf_status kf_pipeline_destroy(kf_pipeline *pl);
/* One call = one processNextFrame: frame 1 becomes the canonical SDF; later frames are
 * integrated into a fresh live SDF, deformed onto the canonical one, warped and fused. */
kf_status kf_pipeline_process_frame(kf_pipeline *pl, const float *depth, kf_frame_stats *stats);
/* Selects which of the reference's per-frame meshes later kf_pipeline_process_frame calls
 * extract (a kf_frame_mesh mask; 0, the default, extracts none), at isovalue iso (the
 * reference's 0, SDF.cpp:410).  On the first frame all three are the canonical SDF's mesh, as
 * in the reference (KillingFusion.cpp:117-118). */
kf_status kf_pipeline_set_meshes(kf_pipeline *pl, int32_t which, double iso);
/* One mesh (a single kf_frame_mesh bit) of the last processed frame, as kf_marching_cubes returns
 * a soup: *n_tris always receives the count; verts may be NULL (size query); KF_E_CAPACITY if
 * cap_tris < count; KF_E_INVALID_ARG if that mesh was not extracted for the frame. */
kf_status kf_pipeline_mesh(kf_pipeline *pl, int32_t which, uint64_t *n_tris, float *verts, uint64_t cap_tris);
kf_volume *kf_pipeline_canonical(kf_pipeline *pl);
kf_volume *kf_pipeline_live(kf_pipeline *pl);
kf_field *kf_pipeline_field(kf_pipeline *pl);

#ifdef __cplusplus
}
#endif
#endif
