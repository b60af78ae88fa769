// kf_pipeline.hip -- the per-frame driver: KillingFusion::processNextFrame
// (KillingFusion/src/KillingFusion.cpp:109-175) over a GPU-resident canonical SDF.
//
// Per frame, in the reference's order (KillingFusion.cpp:130-170):
//   frame 1: canonical <- integrate(frame); field <- 0
//   later:   live <- integrate(frame) into a fresh SDF (computeSDF, :209-231)
//            field <- 0 (UseZeroDisplacementFieldForNextFrame, config.cpp:24)
//            deform(live -> canonical, field)              (:148)
//            live <- warp(live, field)                    (:158, SDF::update)
//            canonical <- fuse(canonical, live)            (:159)
// The reference's three per-frame meshes (getMesh at :157, :160 and :170, returned at
// :171-173) are extracted during the frame when kf_pipeline_set_meshes selects them: the live
// SDF before the warp, the live SDF after it, the canonical SDF after the fuse.  Each is kept
// in a device buffer of the pipeline until the next frame; kf_pipeline_mesh copies one out.
// On the first frame all three are the canonical SDF's mesh (:117-118: the canonical SDF, and
// the canonical SDF fused into an empty one through a zero field, which holds the same values).
// Phase timings use HIP events on the context stream and match the reference's printf columns
// (:127, :167): the fuse column includes the two live meshes, as the reference's timer does
// (:155-161); the canonical mesh is outside the total.
#include "kf_internal.hpp"

#include <cstring>

namespace {

constexpr int NUM_MESHES = 3;  // KF_MESH_LIVE, KF_MESH_LIVE_WARPED, KF_MESH_CANONICAL

struct FrameMesh {
    float *verts = nullptr;  // device, 9 floats per triangle
    size_t cap = 0;          // triangles
    uint64_t count = 0;
    int source = -1;         // slot whose buffer holds this mesh; -1: not extracted this frame
};

int mesh_slot(int32_t which)
{
    return which == KF_MESH_LIVE ? 0 : which == KF_MESH_LIVE_WARPED ? 1 : which == KF_MESH_CANONICAL ? 2 : -1;
}

}  // namespace

struct kf_pipeline {
    kf_context *ctx = nullptr;
    kf_pipeline_desc desc{};
    kf_volume *can = nullptr;
    kf_volume *live = nullptr;
    kf_field *field = nullptr;
    int frame = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t mesh_mask = 0;
    double mesh_iso = 0.0;
    FrameMesh mesh[NUM_MESHES];
};

namespace {

// Extracts v's mesh into slot k of the pipeline's frame meshes.
kf_status extract(kf_pipeline *pl, int k, const kf_volume *v)
{
    FrameMesh &m = pl->mesh[k];
    const kf_status s = kf_mc_to_device(v, pl->mesh_iso, &m.verts, &m.cap, &m.count);
    m.source = s == KF_OK ? k : -1;
    return s;
}

}  // namespace

extern "C" {

kf_status kf_pipeline_destroy(kf_pipeline *pl)
{
    if (!pl)
        return KF_OK;
    kf_volume_destroy(pl->can);
    kf_volume_destroy(pl->live);
    kf_field_destroy(pl->field);
    for (FrameMesh &m : pl->mesh)
        if (m.verts)
            (void)hipFree(m.verts);
    for (int i = 0; i < 4; i++)
        if (pl->ev[i])
            (void)hipEventDestroy(pl->ev[i]);
    delete pl;
    return KF_OK;
}

kf_status kf_pipeline_create(kf_context *ctx, const kf_pipeline_desc *desc, kf_pipeline **out)
{
    KF_CHECK_ARG(ctx && desc && out, "null argument");
    *out = nullptr;
    KF_CHECK_ARG(desc->width > 0 && desc->height > 0, "bad frame size");
    KF_CHECK_ARG(desc->fuse_mode == KF_FUSE_MERGE || desc->fuse_mode == KF_FUSE_NEAREST, "unknown fuse mode");
    kf_pipeline *pl = new kf_pipeline();
    pl->ctx = ctx;
    pl->desc = *desc;
    kf_status s = kf_volume_create(ctx, &desc->volume, &pl->can);
    if (s == KF_OK)
        s = kf_volume_create(ctx, &desc->volume, &pl->live);
    if (s == KF_OK)
        s = kf_field_create(ctx, desc->volume.dims, &pl->field);
    for (int i = 0; i < 4 && s == KF_OK; i++)
        if (hipEventCreate(&pl->ev[i]) != hipSuccess) {
            kf_set_error("hipEventCreate failed");
            s = KF_E_HIP;
        }
    if (s != KF_OK) {
        kf_pipeline_destroy(pl);
        return s;
    }
    *out = pl;
    return KF_OK;
}

kf_status kf_pipeline_process_frame(kf_pipeline *pl, const float *depth, kf_frame_stats *stats)
{
    KF_CHECK_ARG(pl && depth, "null argument");
    KfDeviceGuard dg(pl->ctx->device);
    const kf_pipeline_desc &d = pl->desc;
    hipStream_t st = pl->ctx->stream;
    kf_frame_stats fs;
    std::memset(&fs, 0, sizeof(fs));
    fs.frame_index = pl->frame;
    fs.deform.diverged_index = -1;
    const double zero[3] = {0.0, 0.0, 0.0};
    kf_status s = KF_OK;
    for (FrameMesh &m : pl->mesh)
        m.source = -1;
    const int32_t want = pl->mesh_mask;
    KF_HIP(hipEventRecord(pl->ev[0], st));
    if (pl->frame == 0) {
        // reset + integrate in one pass
        s = kf_integrate_depth_impl(pl->can, depth, d.width, d.height, d.K, d.C2W, d.min_depth, d.max_depth, true);
        if (s == KF_OK)
            s = kf_field_fill(pl->field, zero);
        KF_HIP(hipEventRecord(pl->ev[1], st));
        KF_HIP(hipEventRecord(pl->ev[2], st));
        KF_HIP(hipEventRecord(pl->ev[3], st));
        if (s == KF_OK && want) {
            s = extract(pl, 2, pl->can);
            for (int k = 0; k < NUM_MESHES; k++)
                if (want & (1 << k))
                    pl->mesh[k].source = 2;
            if (!(want & KF_MESH_CANONICAL))
                pl->mesh[2].source = -1;
        }
    } else {
        s = kf_integrate_depth_impl(pl->live, depth, d.width, d.height, d.K, d.C2W, d.min_depth, d.max_depth,
                                    true);
        KF_HIP(hipEventRecord(pl->ev[1], st));
        if (s == KF_OK)
            s = kf_field_fill(pl->field, zero);
        if (s == KF_OK && (d.params.use_data || d.params.use_levelset || d.params.use_killing))
            s = kf_deform_impl(pl->live, pl->can, pl->field, &d.params, &fs.deform, true);  // u was just zeroed
        KF_HIP(hipEventRecord(pl->ev[2], st));
        if (s == KF_OK && (want & KF_MESH_LIVE))
            s = extract(pl, 0, pl->live);
        if (s == KF_OK) {
            // an active-set deform leaves the list of every voxel the warp can change
            bool warped = false;
            s = kf_warp_after_deform(pl->live, pl->field, &warped);
            if (s == KF_OK && !warped)
                s = kf_warp(pl->live, pl->field);
        }
        if (s == KF_OK)
            s = kf_fuse_mode(pl->can, pl->live, d.fuse_mode);
        if (s == KF_OK && (want & KF_MESH_LIVE_WARPED))
            s = extract(pl, 1, pl->live);
        KF_HIP(hipEventRecord(pl->ev[3], st));
        if (s == KF_OK && (want & KF_MESH_CANONICAL))
            s = extract(pl, 2, pl->can);
    }
    KF_HIP(hipEventSynchronize(pl->ev[3]));
    float a = 0.f, b = 0.f, c = 0.f, t = 0.f;
    (void)hipEventElapsedTime(&a, pl->ev[0], pl->ev[1]);
    (void)hipEventElapsedTime(&b, pl->ev[1], pl->ev[2]);
    (void)hipEventElapsedTime(&c, pl->ev[2], pl->ev[3]);
    (void)hipEventElapsedTime(&t, pl->ev[0], pl->ev[3]);
    fs.ms_integrate = a;
    fs.ms_deform = b;
    fs.ms_fuse = c;
    fs.ms_total = t;
    pl->frame += 1;
    if (stats)
        *stats = fs;
    return s;
}

kf_status kf_pipeline_set_meshes(kf_pipeline *pl, int32_t which, double iso)
{
    KF_CHECK_ARG(pl, "null argument");
    KF_CHECK_ARG((which & ~(KF_MESH_LIVE | KF_MESH_LIVE_WARPED | KF_MESH_CANONICAL)) == 0, "unknown mesh bits");
    pl->mesh_mask = which;
    pl->mesh_iso = iso;
    return KF_OK;
}

kf_status kf_pipeline_mesh(kf_pipeline *pl, int32_t which, uint64_t *n_tris, float *verts, uint64_t cap_tris)
{
    KF_CHECK_ARG(pl && n_tris, "null argument");
    const int k = mesh_slot(which);
    KF_CHECK_ARG(k >= 0, "which must be one of KF_MESH_LIVE, KF_MESH_LIVE_WARPED, KF_MESH_CANONICAL");
    KF_CHECK_ARG(pl->mesh[k].source >= 0, "that mesh was not extracted for the last frame (kf_pipeline_set_meshes)");
    const FrameMesh &m = pl->mesh[pl->mesh[k].source];
    *n_tris = m.count;
    if (!verts || m.count == 0)
        return KF_OK;
    if (cap_tris < m.count) {
        kf_set_error("kf_pipeline_mesh: capacity below triangle count");
        return KF_E_CAPACITY;
    }
    KfDeviceGuard dg(pl->ctx->device);
    KF_HIP(hipMemcpyAsync(verts, m.verts, sizeof(float) * 9 * (size_t)m.count, hipMemcpyDeviceToHost, pl->ctx->stream));
    KF_HIP(hipStreamSynchronize(pl->ctx->stream));
    return KF_OK;
}

kf_volume *kf_pipeline_canonical(kf_pipeline *pl) { return pl ? pl->can : nullptr; }
kf_volume *kf_pipeline_live(kf_pipeline *pl) { return pl ? pl->live : nullptr; }
kf_field *kf_pipeline_field(kf_pipeline *pl) { return pl ? pl->field : nullptr; }

}  // extern "C"
