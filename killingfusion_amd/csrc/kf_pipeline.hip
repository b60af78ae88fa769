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
// Meshes are extracted on request with kf_marching_cubes (the reference calls getMesh at
// :157, :160 and :170).  Phase timings use HIP events on the context stream and match the
// reference's printf columns (:127, :167).
#include "kf_internal.hpp"

#include <cstring>

struct kf_pipeline {
    kf_context *ctx = nullptr;
    kf_pipeline_desc desc{};
    kf_volume *can = nullptr;
    kf_volume *live = nullptr;
    kf_field *field = nullptr;
    int frame = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

extern "C" {

kf_status kf_pipeline_destroy(kf_pipeline *pl)
{
    if (!pl)
        return KF_OK;
    kf_volume_destroy(pl->can);
    kf_volume_destroy(pl->live);
    kf_field_destroy(pl->field);
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
    const kf_pipeline_desc &d = pl->desc;
    hipStream_t st = pl->ctx->stream;
    kf_frame_stats fs;
    std::memset(&fs, 0, sizeof(fs));
    fs.frame_index = pl->frame;
    fs.deform.diverged_index = -1;
    const double zero[3] = {0.0, 0.0, 0.0};
    kf_status s = KF_OK;
    KF_HIP(hipEventRecord(pl->ev[0], st));
    if (pl->frame == 0) {
        // reset + integrate in one pass
        s = kf_integrate_depth_impl(pl->can, depth, d.width, d.height, d.K, d.C2W, d.min_depth, d.max_depth, true);
        if (s == KF_OK)
            s = kf_field_fill(pl->field, zero);
        KF_HIP(hipEventRecord(pl->ev[1], st));
        KF_HIP(hipEventRecord(pl->ev[2], st));
        KF_HIP(hipEventRecord(pl->ev[3], st));
    } else {
        s = kf_integrate_depth_impl(pl->live, depth, d.width, d.height, d.K, d.C2W, d.min_depth, d.max_depth,
                                    true);
        KF_HIP(hipEventRecord(pl->ev[1], st));
        if (s == KF_OK)
            s = kf_field_fill(pl->field, zero);
        if (s == KF_OK && (d.params.use_data || d.params.use_levelset || d.params.use_killing))
            s = kf_deform_impl(pl->live, pl->can, pl->field, &d.params, &fs.deform, true);  // u was just zeroed
        KF_HIP(hipEventRecord(pl->ev[2], st));
        if (s == KF_OK) {
            // an active-set deform leaves the list of every voxel the warp can change
            bool warped = false;
            s = kf_warp_after_deform(pl->live, pl->field, &warped);
            if (s == KF_OK && !warped)
                s = kf_warp(pl->live, pl->field);
        }
        if (s == KF_OK)
            s = kf_fuse(pl->can, pl->live);
        KF_HIP(hipEventRecord(pl->ev[3], st));
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

kf_volume *kf_pipeline_canonical(kf_pipeline *pl) { return pl ? pl->can : nullptr; }
kf_volume *kf_pipeline_live(kf_pipeline *pl) { return pl ? pl->live : nullptr; }
kf_field *kf_pipeline_field(kf_pipeline *pl) { return pl ? pl->field : nullptr; }

}  // extern "C"
