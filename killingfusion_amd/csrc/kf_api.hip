// kf_api.hip -- the C ABI of libkf (include/kf_api.h): handles, transfers and the
// optimiser driver.  Kernels live in k_*.hip.
#include "kf_reduce.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstring>
#include <string>

kf_status kf_mc_init_tables();
kf_status kf_noalloc_init();

namespace {

thread_local std::string g_last_error;

constexpr int STATE_BLOCK = 64;

__global__ void k_state_reset(DeformState *st, double stop)
{
    if (threadIdx.x == 0) {
        st->done = 0;
        st->iters = 0;
        st->run[0] = 1;
        st->run[1] = 1;
        st->diverged = LLONG_MAX;
        st->div_slot[0] = LLONG_MAX;
        st->div_slot[1] = LLONG_MAX;
        st->stop = stop;
        st->last_max = 0.0;
        st->energy = 0.0;
        st->last_active = 0ull;
        st->listed = 0u;
        st->nonfinite = 0u;
        st->list_n = 0u;
        st->list_energy = 0.0;
        st->list_bad = 0u;
    }
}

using DeviceGuard = KfDeviceGuard;

bool valid_dims(const int32_t d[3])
{
    if (!d)
        return false;
    for (int a = 0; a < 3; a++)
        if (d[a] <= 0 || d[a] > (1 << 15))
            return false;
    return (int64_t)d[0] * d[1] * d[2] <= (int64_t)1 << 33;
}

bool valid_params_impl(const kf_params *p)
{
    // the trust region is the reference formula's (KillingFusion.cpp:288-314); the adjoint
    // method has no counterpart of it
    return p && p->voxel_size > 0.0 && p->max_iters >= 0 && p->delta_size > 0.0 &&
           (p->method == KF_METHOD_REFERENCE || p->method == KF_METHOD_ADJOINT) &&
           (p->sweep == KF_SWEEP_AUTO || p->sweep == KF_SWEEP_DENSE || p->sweep == KF_SWEEP_ACTIVE) &&
           (p->use_trust == 0 || (p->use_trust == 1 && p->method == KF_METHOD_REFERENCE)) && p->reserved == 0;
}

// Gauss-Jordan inverse with partial pivoting (row-major 4x4).
bool invert4(const double m[16], double out[16])
{
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 8; c++)
            a[r][c] = c < 4 ? m[r * 4 + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c]))
                piv = r;
        if (a[piv][c] == 0.0)
            return false;
        if (piv != c)
            for (int k = 0; k < 8; k++)
                std::swap(a[c][k], a[piv][k]);
        const double d = a[c][c];
        for (int k = 0; k < 8; k++)
            a[c][k] /= d;
        for (int r = 0; r < 4; r++) {
            if (r == c || a[r][c] == 0.0)
                continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; k++)
                a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            out[r * 4 + c] = a[r][c + 4];
    return true;
}

// Allocates the field's two mask buffers on first use: each holds the stencil mask and the
// nonzero mask it is dilated from (kf_launch_adj_mask).
kf_status ensure_masks(kf_field *f)
{
    const size_t words = kf_adj_mask_words(f->dims);
    for (int b = 0; b < 2; b++)
        if (!f->masks[b])
            KF_HIP(hipMalloc((void **)&f->masks[b], 2 * sizeof(unsigned int) * words));
    return KF_OK;
}

kf_status read_state(kf_context *ctx)
{
    KF_HIP(hipMemcpyAsync(ctx->state_host, ctx->state, sizeof(DeformState), hipMemcpyDeviceToHost, ctx->stream));
    KF_HIP(hipStreamSynchronize(ctx->stream));
    return KF_OK;
}

}  // namespace

void kf_set_error(const std::string &msg) { g_last_error = msg; }

uint64_t kf_next_version()
{
    static std::atomic<uint64_t> counter{0};
    return counter.fetch_add(1, std::memory_order_relaxed) + 1;
}

bool kf_valid_params(const kf_params *p) { return valid_params_impl(p); }

kf_status kf_ensure_partials(kf_context *ctx, size_t nblocks)
{
    if (nblocks <= ctx->partials_cap)
        return KF_OK;
    if (ctx->partials) {
        KF_HIP(hipStreamSynchronize(ctx->stream));
        KF_HIP(hipFree(ctx->partials));
        ctx->partials = nullptr;
        ctx->partials_cap = 0;
    }
    KF_HIP(hipMalloc((void **)&ctx->partials, 2 * sizeof(BlockPartial) * nblocks));
    ctx->partials_cap = nblocks;
    return KF_OK;
}

kf_status kf_ensure_need_words(kf_context *ctx, size_t words)
{
    if (words <= ctx->need_cap)
        return KF_OK;
    if (ctx->need_words) {
        KF_HIP(hipStreamSynchronize(ctx->stream));
        KF_HIP(hipFree(ctx->need_words));
        ctx->need_words = nullptr;
        ctx->need_cap = 0;
    }
    KF_HIP(hipMalloc((void **)&ctx->need_words, sizeof(unsigned long long) * words));
    ctx->need_cap = words;
    return KF_OK;
}

namespace {

constexpr int FIN_BLOCK = 1024;

// Ends one deform iteration: folds one step launch's partials in index order (so the FP64
// energy sum is reproducible), records max |du| and the active count, advances the
// iteration counter and raises `done` when the stop rule fires or a voxel diverged.
__global__ __launch_bounds__(FIN_BLOCK) void k_finalize(DeformState *st, const BlockPartial *parts, unsigned int nb,
                                                        int parity)
{
    if (st->done)
        return;
    kf_fold_launch<FIN_BLOCK>(st, parts, nb, parity);
}

}  // namespace

kf_status kf_launch_finalize(kf_context *ctx, unsigned int nb, int parity)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_BLOCK), 0, ctx->stream, ctx->state,
                       (const BlockPartial *)kf_partials(ctx, parity), nb, parity);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

kf_status kf_state_reset(kf_context *ctx, double stop)
{
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(STATE_BLOCK), 0, ctx->stream, ctx->state, stop);
    KF_LAUNCH_CHECK();
    return KF_OK;
}

extern "C" {

const char *kf_status_string(kf_status s)
{
    switch (s) {
    case KF_OK: return "ok";
    case KF_E_INVALID_ARG: return "invalid argument";
    case KF_E_DIVERGED: return "displacement field diverged";
    case KF_E_OOM: return "out of device memory";
    case KF_E_HIP: return "HIP runtime error";
    case KF_E_NO_DEVICE: return "no usable gfx950 device";
    case KF_E_CAPACITY: return "output capacity too small";
    }
    return "unknown status";
}

const char *kf_last_error(void) { return g_last_error.c_str(); }

int kf_api_version(void) { return KF_API_VERSION; }

kf_status kf_host_alloc(size_t bytes, void **out)
{
    KF_CHECK_ARG(out && bytes > 0, "null out or zero size");
    *out = nullptr;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        kf_set_error("page-locked host allocation failed");
        return KF_E_OOM;
    }
    return KF_OK;
}

kf_status kf_host_free(void *p)
{
    if (p)
        KF_HIP(hipHostFree(p));
    return KF_OK;
}

kf_status kf_params_default(kf_params *p)
{
    KF_CHECK_ARG(p, "params is null");
    std::memset(p, 0, sizeof(*p));
    p->voxel_size = 0.012;                 // config.cpp:4
    p->max_surface = 0.012 * 4;            // config.cpp:6
    p->clip_distance = 0.012 * 4;          // config.cpp:5
    p->epsilon = 0.00001;                  // config.cpp:57
    p->alpha = 0.025;                      // config.cpp:45
    p->omega_killing = 0.04;               // config.cpp:48
    p->gamma_killing = 0.0;                // config.cpp:51
    p->omega_levelset = 0.1;               // config.cpp:54
    p->delta_size = 0.05;                  // config.cpp:29
    p->killing_clamp = 0.1;                // DisplacementField.cpp:342
    p->use_data = p->use_levelset = p->use_killing = 1;  // config.cpp:23
    p->max_iters = 1024;                   // config.cpp:42
    p->stop_max_update = 0.0;
    p->method = KF_METHOD_REFERENCE;
    p->use_trust = 0;                      // config.cpp:27
    return KF_OK;
}

kf_status kf_context_create(int device, void *stream, kf_context **out)
{
    KF_CHECK_ARG(out, "out is null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        kf_set_error("no HIP device visible");
        return KF_E_NO_DEVICE;
    }
    KF_CHECK_ARG(device >= 0 && device < count, "device index out of range");
    hipDeviceProp_t prop;
    KF_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        kf_set_error(std::string("device is ") + prop.gcnArchName + ", libkf is built for gfx950");
        return KF_E_NO_DEVICE;
    }
    DeviceGuard dg(device);
    kf_context *ctx = new kf_context();
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount;
    kf_status s = KF_OK;
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            kf_set_error("hipStreamCreate failed");
            delete ctx;
            return KF_E_HIP;
        }
        ctx->own_stream = true;
    }
    if (hipMalloc((void **)&ctx->state, sizeof(DeformState)) != hipSuccess ||
        hipHostMalloc((void **)&ctx->state_host, sizeof(DeformState), hipHostMallocDefault) != hipSuccess) {
        kf_set_error("context allocation failed");
        kf_context_destroy(ctx);
        return KF_E_OOM;
    }
    s = kf_mc_init_tables();
    if (s == KF_OK)
        s = kf_noalloc_init();
    if (s == KF_OK)
        s = kf_state_reset(ctx, 0.0);
    if (s != KF_OK) {
        kf_context_destroy(ctx);
        return s;
    }
    *out = ctx;
    return KF_OK;
}

kf_status kf_context_destroy(kf_context *ctx)
{
    if (!ctx)
        return KF_OK;
    DeviceGuard dg(ctx->device);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    if (ctx->state)
        (void)hipFree(ctx->state);
    if (ctx->state_host)
        (void)hipHostFree(ctx->state_host);
    if (ctx->partials)
        (void)hipFree(ctx->partials);
    if (ctx->need_words)
        (void)hipFree(ctx->need_words);
    kf_active_release(ctx);
    kf_ref_release(ctx);
    kf_batch_release(ctx);
    kf_mc_release(ctx);
    if (ctx->own_stream && ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return KF_OK;
}

kf_status kf_context_synchronize(kf_context *ctx)
{
    KF_CHECK_ARG(ctx, "context is null");
    DeviceGuard dg(ctx->device);
    KF_HIP(hipStreamSynchronize(ctx->stream));
    return KF_OK;
}

void *kf_context_stream(kf_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

kf_status kf_context_set_loop_capture(kf_context *ctx, int32_t enable)
{
    KF_CHECK_ARG(ctx, "context is null");
    ctx->capture_loops = enable != 0;
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Volumes
// ---------------------------------------------------------------------------

kf_status kf_dims_from_bounds(const double min_corner[3], const double max_corner[3], double voxel, int32_t dims_out[3])
{
    KF_CHECK_ARG(min_corner && max_corner && dims_out && voxel > 0.0, "bad bounds");
    for (int a = 0; a < 3; a++) {
        const double c = (max_corner[a] - min_corner[a]) / voxel;
        KF_CHECK_ARG(c >= 0.0 && c < (double)(1 << 15), "bounds give an invalid grid");
        dims_out[a] = (int32_t)c + 1;  // SDF.cpp:122
    }
    return KF_OK;
}

kf_status kf_volume_create(kf_context *ctx, const kf_volume_desc *d, kf_volume **out)
{
    KF_CHECK_ARG(ctx && d && out, "null argument");
    *out = nullptr;
    KF_CHECK_ARG(valid_dims(d->dims), "dims must be positive and at most 32768 per axis");
    KF_CHECK_ARG(d->voxel_size > 0.0, "voxel_size must be positive");
    KF_CHECK_ARG(d->dtype == KF_F32 || d->dtype == KF_BF16 || d->dtype == KF_F16, "unknown dtype");
    KF_CHECK_ARG(d->max_surface + d->epsilon > 0.0, "max_surface + epsilon must be positive");
    DeviceGuard dg(ctx->device);
    kf_volume *v = new kf_volume();
    v->ctx = ctx;
    for (int a = 0; a < 3; a++) {
        v->dims[a] = d->dims[a];
        v->min_corner[a] = d->min_corner[a];
    }
    v->n = (int64_t)d->dims[0] * d->dims[1] * d->dims[2];
    v->voxel = d->voxel_size;
    v->clip = d->clip_distance;
    v->max_surface = d->max_surface;
    v->epsilon = d->epsilon;
    v->dtype = (kf_dtype)d->dtype;
    if (hipMalloc(&v->phi, v->elem_size() * (size_t)v->n) != hipSuccess ||
        hipMalloc((void **)&v->w, sizeof(int32_t) * (size_t)v->n) != hipSuccess) {
        kf_set_error("volume allocation failed");
        kf_volume_destroy(v);
        return KF_E_OOM;
    }
    kf_status s = kf_launch_fill_volume(v);
    if (s != KF_OK) {
        kf_volume_destroy(v);
        return s;
    }
    *out = v;
    return KF_OK;
}

kf_status kf_volume_destroy(kf_volume *v)
{
    if (!v)
        return KF_OK;
    DeviceGuard dg(v->ctx->device);
    (void)hipStreamSynchronize(v->ctx->stream);
    if (v->phi)
        (void)hipFree(v->phi);
    if (v->w)
        (void)hipFree(v->w);
    delete v;
    return KF_OK;
}

kf_status kf_volume_describe(const kf_volume *v, kf_volume_desc *out)
{
    KF_CHECK_ARG(v && out, "null argument");
    std::memset(out, 0, sizeof(*out));
    for (int a = 0; a < 3; a++) {
        out->dims[a] = v->dims[a];
        out->min_corner[a] = v->min_corner[a];
    }
    out->voxel_size = v->voxel;
    out->clip_distance = v->clip;
    out->max_surface = v->max_surface;
    out->epsilon = v->epsilon;
    out->dtype = v->dtype;
    return KF_OK;
}

kf_status kf_volume_reset(kf_volume *v)
{
    KF_CHECK_ARG(v, "volume is null");
    DeviceGuard dg(v->ctx->device);
    return kf_launch_fill_volume(v);
}

kf_status kf_volume_upload(kf_volume *v, const float *phi, const int32_t *weight)
{
    KF_CHECK_ARG(v && phi, "null argument");
    DeviceGuard dg(v->ctx->device);
    hipStream_t s = v->ctx->stream;
    float *tmp = nullptr;
    KF_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * (size_t)v->n, s));
    hipError_t e = hipMemcpyAsync(tmp, phi, sizeof(float) * (size_t)v->n, hipMemcpyHostToDevice, s);
    kf_status st = KF_OK;
    if (e == hipSuccess)
        st = kf_launch_upload_volume(v, tmp, nullptr);
    if (e == hipSuccess && st == KF_OK) {
        if (weight)
            e = hipMemcpyAsync(v->w, weight, sizeof(int32_t) * (size_t)v->n, hipMemcpyHostToDevice, s);
        else
            e = hipMemsetAsync(v->w, 0, sizeof(int32_t) * (size_t)v->n, s);
    }
    (void)hipFreeAsync(tmp, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        kf_set_error(std::string("volume upload: ") + hipGetErrorString(e));
        return KF_E_HIP;
    }
    return st;
}

kf_status kf_volume_download(const kf_volume *v, float *phi, int32_t *weight)
{
    KF_CHECK_ARG(v, "volume is null");
    DeviceGuard dg(v->ctx->device);
    hipStream_t s = v->ctx->stream;
    kf_status st = KF_OK;
    hipError_t e = hipSuccess;
    if (phi) {
        float *tmp = nullptr;
        KF_HIP(hipMallocAsync((void **)&tmp, sizeof(float) * (size_t)v->n, s));
        st = kf_launch_download_volume(v, tmp);
        if (st == KF_OK)
            e = hipMemcpyAsync(phi, tmp, sizeof(float) * (size_t)v->n, hipMemcpyDeviceToHost, s);
        (void)hipFreeAsync(tmp, s);
    }
    if (weight && e == hipSuccess && st == KF_OK)
        e = hipMemcpyAsync(weight, v->w, sizeof(int32_t) * (size_t)v->n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        kf_set_error(std::string("volume download: ") + hipGetErrorString(e));
        return KF_E_HIP;
    }
    return st;
}

kf_status kf_volume_copy(kf_volume *dst, const kf_volume *src)
{
    KF_CHECK_ARG(dst && src, "null argument");
    KF_CHECK_ARG(same_dims(dst->dims, src->dims) && dst->dtype == src->dtype, "volumes differ in dims or dtype");
    KF_CHECK_ARG(dst->ctx == src->ctx, "handles belong to different contexts");
    DeviceGuard dg(dst->ctx->device);
    hipStream_t s = dst->ctx->stream;
    kf_touch(dst);
    KF_HIP(hipMemcpyAsync(dst->phi, src->phi, dst->elem_size() * (size_t)dst->n, hipMemcpyDeviceToDevice, s));
    KF_HIP(hipMemcpyAsync(dst->w, src->w, sizeof(int32_t) * (size_t)dst->n, hipMemcpyDeviceToDevice, s));
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Displacement fields
// ---------------------------------------------------------------------------

kf_status kf_field_create(kf_context *ctx, const int32_t dims[3], kf_field **out)
{
    KF_CHECK_ARG(ctx && out, "null argument");
    *out = nullptr;
    KF_CHECK_ARG(valid_dims(dims), "dims must be positive and at most 32768 per axis");
    DeviceGuard dg(ctx->device);
    kf_field *f = new kf_field();
    f->ctx = ctx;
    for (int a = 0; a < 3; a++)
        f->dims[a] = dims[a];
    f->n = (int64_t)dims[0] * dims[1] * dims[2];
    // Both buffers of the Jacobi pair come from one allocation.  Allocated apart, the second
    // buffer's physical placement decided the dense step's speed: 6 of 32 placements ran the
    // 256^3 step 25% slower (58 -> 72-74 us per launch); in one allocation 80 of 80 ran at
    // 58-59 us.
    if (hipMalloc((void **)&f->buf[0], sizeof(float) * 6 * (size_t)f->n) != hipSuccess) {
        f->buf[0] = nullptr;
        kf_set_error("field allocation failed");
        kf_field_destroy(f);
        return KF_E_OOM;
    }
    f->buf[1] = f->buf[0] + 3 * (size_t)f->n;
    const double zero[3] = {0.0, 0.0, 0.0};
    kf_status s = kf_launch_field_fill(f, zero);
    if (s != KF_OK) {
        kf_field_destroy(f);
        return s;
    }
    *out = f;
    return KF_OK;
}

kf_status kf_field_destroy(kf_field *f)
{
    if (!f)
        return KF_OK;
    DeviceGuard dg(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    if (f->buf[0])
        (void)hipFree(f->buf[0]);  // buf[1] lies in the same allocation
    for (int b = 0; b < 2; b++) {
        if (f->masks[b])
            (void)hipFree(f->masks[b]);
    }
    delete f;
    return KF_OK;
}

kf_status kf_field_fill(kf_field *f, const double value[3])
{
    KF_CHECK_ARG(f && value, "null argument");
    DeviceGuard dg(f->ctx->device);
    return kf_launch_field_fill(f, value);
}

kf_status kf_field_upload(kf_field *f, const float *soa)
{
    KF_CHECK_ARG(f && soa, "null argument");
    DeviceGuard dg(f->ctx->device);
    KF_HIP(hipMemcpyAsync(f->u(), soa, sizeof(float) * 3 * (size_t)f->n, hipMemcpyHostToDevice, f->ctx->stream));
    KF_HIP(hipStreamSynchronize(f->ctx->stream));
    return KF_OK;
}

kf_status kf_field_download(const kf_field *f, float *soa)
{
    KF_CHECK_ARG(f && soa, "null argument");
    DeviceGuard dg(f->ctx->device);
    KF_HIP(hipMemcpyAsync(soa, f->u(), sizeof(float) * 3 * (size_t)f->n, hipMemcpyDeviceToHost, f->ctx->stream));
    KF_HIP(hipStreamSynchronize(f->ctx->stream));
    return KF_OK;
}

kf_status kf_field_add(kf_field *dst, const kf_field *src)
{
    KF_CHECK_ARG(dst && src, "null argument");
    KF_CHECK_ARG(same_dims(dst->dims, src->dims), "fields differ in dims");
    DeviceGuard dg(dst->ctx->device);
    return kf_launch_field_add(dst, src);
}

// ---------------------------------------------------------------------------
// Per-frame kernels
// ---------------------------------------------------------------------------

kf_status kf_integrate_depth(kf_volume *v, const float *depth, int32_t width, int32_t height, const double K[9],
                             const double C2W[16], double min_depth, double max_depth)
{
    return kf_integrate_depth_impl(v, depth, width, height, K, C2W, min_depth, max_depth, false);
}

}  // extern "C"

kf_status kf_integrate_depth_impl(kf_volume *v, const float *depth, int32_t width, int32_t height, const double K[9],
                                  const double C2W[16], double min_depth, double max_depth, bool fresh)
{
    KF_CHECK_ARG(v && depth && K && C2W, "null argument");
    KF_CHECK_ARG(width > 0 && height > 0 && (int64_t)width * height <= ((int64_t)1 << 28), "bad frame size");
    double W2C[16];
    KF_CHECK_ARG(invert4(C2W, W2C), "camera pose is singular");
    DeviceGuard dg(v->ctx->device);
    hipStream_t s = v->ctx->stream;
    float *dd = nullptr;
    const size_t bytes = sizeof(float) * (size_t)width * (size_t)height;
    KF_HIP(hipMallocAsync((void **)&dd, bytes, s));
    hipError_t e = hipMemcpyAsync(dd, depth, bytes, hipMemcpyHostToDevice, s);
    kf_status st = KF_OK;
    if (e == hipSuccess)
        st = kf_launch_integrate(v, dd, width, height, K, W2C, min_depth, max_depth, fresh);
    (void)hipFreeAsync(dd, s);
    if (e != hipSuccess) {
        kf_set_error(std::string("depth upload: ") + hipGetErrorString(e));
        return KF_E_HIP;
    }
    return st;
}


namespace {

// One deform call.  predict: size the list launches from the list length of the previous call
// on a volume of this size instead of reading the length back first.  *rerun: the list was
// longer than predicted (or, for the reference method, met a non-finite live value), every
// list kernel did nothing and the field is unchanged, so the call must run again unpredicted.
kf_status deform_pass(const kf_volume *live, const kf_volume *can, kf_field *u, const kf_params *p,
                      kf_deform_stats *stats, bool u_zero, bool predict, bool *rerun)
{
    *rerun = false;
    kf_context *ctx = live->ctx;
    kf_active_ws &a = ctx->active;
    kf_status s = kf_state_reset(ctx, p->stop_max_update);
    if (s != KF_OK)
        return s;
    a.warp_vol = nullptr;
    const int cur0 = u->cur;
    const bool check = p->stop_max_update > 0.0;
    const bool adj = p->method == KF_METHOD_ADJOINT;
    // Active-set sweep (k_deform_active.hip): count the voxels that can move and list them.  The
    // adjoint method steps only them, in compact arrays, and writes the result back into the
    // field's buffer; the reference method steps only them in the field's own two buffers,
    // which both hold +0 everywhere else (the other buffer starts as a copy of the field).
    bool listed = false, ref_listed = false;
    unsigned int slots = 0, expect = 0;  // list launch slots; the length they are dealt for
    if (p->max_iters > 0 && p->sweep != KF_SWEEP_DENSE && kf_active_supported(live)) {
        // Both sweeps give the same field bit for bit, so which one runs is a question of speed
        // only, and a predicted list length decides it as well as the read-back one.
        bool predicted = predict && a.hint_n == live->n &&
                         (p->sweep == KF_SWEEP_ACTIVE || (int64_t)a.hint_listed * 8 <= live->n);
        if (predicted) {
            // never more slots than a list of every voxel needs
            const unsigned int most = kf_list_slots((unsigned int)std::min<int64_t>(live->n, KF_LIST_MAX), false);
            slots = kf_list_slots(a.hint_listed, true);
            slots = slots < most ? slots : most;
            expect = kf_list_slots(a.hint_listed, false);
            expect = expect < slots ? expect : slots;
            // A list that drifts a little from call to call (a frame sequence) keeps the launch
            // shape of the last predicted call while it still fits with some room, so the
            // captured loop of that shape keeps being replayed.
            const unsigned int room = a.hint_listed + a.hint_listed / 32;
            if (a.shape_n == live->n && room <= a.shape_slots && a.shape_expect <= a.shape_slots &&
                expect + 2048 >= a.shape_expect && a.shape_expect + 2048 >= expect) {
                slots = a.shape_slots;
                expect = a.shape_expect;
            }
            a.shape_n = live->n;
            a.shape_slots = slots;
            a.shape_expect = expect;
        }
        predicted = predicted && slots > 0;
        // The reference method's band test passes a NaN warped value, and with u_j = +0 next to
        // a non-finite live value the warped value is NaN (inf * 0): such a voxel moves though
        // phi_live[j] is outside the band.  The list covers every voxel that can move only when
        // every live value is finite, so the reference method steps densely otherwise.
        s = kf_active_count(live, u->buf[cur0], p, u_zero, predicted ? slots : 0xffffffffu, !adj);
        if (s != KF_OK)
            return s;
        bool use = predicted;
        if (!predicted) {
            s = read_state(ctx);
            if (s != KF_OK)
                return s;
            const unsigned int count = ctx->state_host->listed;
            const bool finite_ok = adj || !ctx->state_host->nonfinite;
            slots = expect = kf_list_slots(count, false);
            use = (p->sweep == KF_SWEEP_ACTIVE || (int64_t)count * 8 <= live->n) && finite_ok && slots > 0;
            a.hint_n = finite_ok ? live->n : 0;
            a.hint_listed = count;
        }
        // room for the launch shape the next call predicts from this list
        const unsigned int most = kf_list_slots((unsigned int)std::min<int64_t>(live->n, KF_LIST_MAX), false);
        const unsigned int ahead = kf_list_slots(ctx->state_host->listed, true);
        const unsigned int reserve = predicted ? 0u : (ahead < most ? ahead : most);
        if (use && (s = kf_active_prepare(live, u->buf[cur0], slots, adj, reserve)) != KF_OK)
            return s;
        listed = use && adj;
        ref_listed = use && !adj;
        if (ref_listed)
            KF_HIP(hipMemcpyAsync(u->buf[cur0 ^ 1], u->buf[cur0], sizeof(float) * 3 * (size_t)live->n,
                                  hipMemcpyDeviceToDevice, ctx->stream));
    }
    // A list step is a few microseconds; check the stop rule less often so the host stays ahead.
    const int CHUNK = listed ? 64 : ref_listed ? 32 : 16;
    // Launch k reads buffer cur0 ^ (k & 1) and writes the other one.  An adjoint launch k folds
    // launch k - 1 in its first block, so launch k itself may run before the stop rule (or a
    // divergence) of k - 1 is known: it only writes the buffer that is not the result, and
    // launch k + 1 sees `done` and does nothing.
    int launched = 0;
    unsigned int nb = 0;
    if (listed && !check && p->max_iters > 0) {
        // the whole loop and its last fold, as one captured graph when this launch shape recurs
        s = kf_active_loop(live, can, p, slots, expect);
        if (s != KF_OK)
            return s;
        launched = -1;  // folded already
    }
    for (int k = 0; launched >= 0 && k < p->max_iters; k++) {
        float *in = u->buf[cur0 ^ (k & 1)];
        float *out = u->buf[cur0 ^ ((k + 1) & 1)];
        s = listed ? kf_launch_adj_list(live, can, p, k, slots, expect, &nb)
            : adj  ? kf_launch_adj_step(live, can, in, out, p, KF_MODE_STEP, nullptr, nullptr, k, &nb)
                   : kf_launch_ref_step(live, can, in, out, p, KF_MODE_STEP, k, ref_listed ? a.list : nullptr,
                                        slots);
        if (s != KF_OK)
            return s;
        launched = k + 1;
        if (check && (k + 1) % CHUNK == 0 && k + 1 < p->max_iters) {
            s = read_state(ctx);
            if (s != KF_OK)
                return s;
            if (ctx->state_host->done)
                break;
        }
    }
    if (adj && launched > 0) {
        s = kf_launch_finalize(ctx, nb, (launched - 1) & 1);
        if (s != KF_OK)
            return s;
    }
    // The step kernels skip the energy of voxels that do not move; one forward pass on the
    // final field gives the energy the optimiser reached.  With a list, the result's compact
    // set (iters & 1) is chosen on the device, so the field, the energy and the state need one
    // readback; the energy of a diverged field is computed and ignored.
    if (listed) {
        s = kf_active_scatter(live, u->buf[cur0], slots);
        if (s == KF_OK)
            s = kf_active_energy(live, can, p, slots, expect);
        if (s != KF_OK)
            return s;
    }
    s = read_state(ctx);
    if (s != KF_OK)
        return s;
    const DeformState h = *ctx->state_host;
    if ((listed || ref_listed) && h.list_bad) {
        a.hint_n = 0;
        *rerun = true;
        return KF_OK;
    }
    if (listed || ref_listed) {
        a.hint_n = live->n;
        a.hint_listed = h.listed;
    }
    if (!listed)
        u->cur = cur0 ^ (int)(h.iters & 1u);
    double energy = 0.0;
    if (p->method == KF_METHOD_ADJOINT && h.diverged == LLONG_MAX) {
        if (listed) {
            energy = h.list_energy;
        } else {
            s = kf_energy_forward(live, can, u, p, &energy, nullptr);
            if (s != KF_OK)
                return s;
        }
    }
    if ((listed || ref_listed) && !h.nonfinite) {
        // The list covers every voxel a warp of live by u can change (see k_deform_active.hip).
        a.warp_vol = live;
        a.warp_field = u;
        a.warp_count = h.listed;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iters = (int32_t)h.iters;
        stats->diverged_index = h.diverged == LLONG_MAX ? -1 : (int64_t)h.diverged;
        stats->max_update = h.last_max;
        stats->energy = energy;
        stats->active_voxels = h.last_active;
        stats->sweep = listed || ref_listed ? KF_SWEEP_ACTIVE : KF_SWEEP_DENSE;
        stats->listed_voxels = h.listed;  // 0 when nothing was counted
    }
    if (h.diverged != LLONG_MAX) {
        kf_set_error("deformation field diverged at voxel " + std::to_string((long long)h.diverged));
        return KF_E_DIVERGED;
    }
    return KF_OK;
}

}  // namespace

kf_status kf_deform_impl(const kf_volume *live, const kf_volume *can, kf_field *u, const kf_params *p,
                         kf_deform_stats *stats, bool u_zero)
{
    kf_status s = kf_check_triplet(live, can, u);
    if (s != KF_OK)
        return s;
    KF_CHECK_ARG(kf_valid_params(p), "invalid params");
    DeviceGuard dg(live->ctx->device);
    bool rerun = false;
    s = deform_pass(live, can, u, p, stats, u_zero, true, &rerun);
    if (s == KF_OK && rerun)
        s = deform_pass(live, can, u, p, stats, u_zero, false, &rerun);
    return s;
}

kf_status kf_check_triplet(const kf_volume *live, const kf_volume *can, const kf_field *u)
{
    KF_CHECK_ARG(live && can && u, "null argument");
    KF_CHECK_ARG(same_dims(live->dims, can->dims) && same_dims(live->dims, u->dims),
                 "live, canonical and field dims differ");
    KF_CHECK_ARG(live->ctx == can->ctx && live->ctx == u->ctx, "handles belong to different contexts");
    KF_CHECK_ARG(live->dtype == can->dtype, "live and canonical volumes must share a dtype");
    return KF_OK;
}

extern "C" {

kf_status kf_energy_gradient_ref(const kf_volume *live, const kf_volume *can, const kf_field *u, const kf_params *p,
                                 kf_field *grad_out)
{
    kf_status s = kf_check_triplet(live, can, u);
    if (s != KF_OK)
        return s;
    KF_CHECK_ARG(kf_valid_params(p), "invalid params");
    KF_CHECK_ARG(grad_out && grad_out != u && same_dims(grad_out->dims, u->dims), "grad_out must be a distinct field of the same dims");
    DeviceGuard dg(live->ctx->device);
    return kf_launch_ref_step(live, can, u->u(), grad_out->u(), p, KF_MODE_GRAD);
}

kf_status kf_energy_forward(const kf_volume *live, const kf_volume *can, const kf_field *u, const kf_params *p,
                            double *energy_out, float *energy_per_voxel)
{
    kf_status s = kf_check_triplet(live, can, u);
    if (s != KF_OK)
        return s;
    KF_CHECK_ARG(kf_valid_params(p) && energy_out, "invalid params or null energy_out");
    kf_context *ctx = live->ctx;
    DeviceGuard dg(ctx->device);
    float *ev = nullptr;
    if (energy_per_voxel)
        KF_HIP(hipMallocAsync((void **)&ev, sizeof(float) * (size_t)live->n, ctx->stream));
    s = kf_state_reset(ctx, 0.0);
    const unsigned int *mask = nullptr;
    if (s == KF_OK && kf_adj_mask_words(u->dims) > 0) {
        kf_field *uf = const_cast<kf_field *>(u);  // the masks are a cache owned by the field
        s = ensure_masks(uf);
        if (s == KF_OK)
            s = kf_launch_adj_mask(uf, uf->u(), uf->masks[uf->cur]);
        mask = uf->masks[uf->cur];
    }
    if (s == KF_OK)
        s = kf_launch_adj_step(live, can, u->u(), nullptr, p, KF_MODE_ENERGY, ev, mask);
    if (s == KF_OK && ev) {
        hipError_t e = hipMemcpyAsync(energy_per_voxel, ev, sizeof(float) * (size_t)live->n, hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e != hipSuccess) {
            kf_set_error(std::string("energy download: ") + hipGetErrorString(e));
            s = KF_E_HIP;
        }
    }
    if (ev)
        (void)hipFreeAsync(ev, ctx->stream);
    if (s == KF_OK)
        s = read_state(ctx);
    if (s == KF_OK)
        *energy_out = ctx->state_host->energy;
    return s;
}

kf_status kf_energy_backward(const kf_volume *live, const kf_volume *can, const kf_field *u, const kf_params *p,
                             kf_field *grad_out)
{
    kf_status s = kf_check_triplet(live, can, u);
    if (s != KF_OK)
        return s;
    KF_CHECK_ARG(kf_valid_params(p), "invalid params");
    KF_CHECK_ARG(grad_out && grad_out != u && same_dims(grad_out->dims, u->dims), "grad_out must be a distinct field of the same dims");
    DeviceGuard dg(live->ctx->device);
    const unsigned int *mask = nullptr;
    if (kf_adj_mask_words(u->dims) > 0) {
        kf_field *uf = const_cast<kf_field *>(u);  // the masks are a cache owned by the field
        s = ensure_masks(uf);
        if (s == KF_OK)
            s = kf_launch_adj_mask(uf, uf->u(), uf->masks[uf->cur]);
        if (s != KF_OK)
            return s;
        mask = uf->masks[uf->cur];
    }
    return kf_launch_adj_step(live, can, u->u(), grad_out->u(), p, KF_MODE_GRAD, nullptr, mask);
}

kf_status kf_deform(const kf_volume *live, const kf_volume *can, kf_field *u, const kf_params *p,
                    kf_deform_stats *stats)
{
    return kf_deform_impl(live, can, u, p, stats, false);
}

kf_status kf_warp(kf_volume *v, const kf_field *u)
{
    KF_CHECK_ARG(v && u, "null argument");
    KF_CHECK_ARG(same_dims(v->dims, u->dims), "volume and field dims differ");
    DeviceGuard dg(v->ctx->device);
    // Only voxels with u != 0 can change (k_moved_mark); when they are few, warp those in
    // place.
    bool sparse = false;
    const kf_status ss = kf_warp_sparse(v, u, &sparse);
    if (ss != KF_OK || sparse)
        return ss;
    hipStream_t s = v->ctx->stream;
    void *np = nullptr;
    int32_t *nw = nullptr;
    KF_HIP(hipMallocAsync(&np, v->elem_size() * (size_t)v->n, s));
    hipError_t e = hipMallocAsync((void **)&nw, sizeof(int32_t) * (size_t)v->n, s);
    if (e != hipSuccess) {
        (void)hipFreeAsync(np, s);
        kf_set_error("warp scratch allocation failed");
        return KF_E_OOM;
    }
    kf_status st = kf_launch_warp(v, u, np, nw);
    if (st != KF_OK) {
        (void)hipFreeAsync(np, s);
        (void)hipFreeAsync(nw, s);
        return st;
    }
    // The reference swaps the new grids in (SDF.cpp:326-327); so does this.
    (void)hipFreeAsync(v->phi, s);
    (void)hipFreeAsync(v->w, s);
    v->phi = np;
    v->w = nw;
    kf_touch(v);
    return KF_OK;
}

kf_status kf_fuse_mode(kf_volume *can, const kf_volume *live, int32_t mode)
{
    KF_CHECK_ARG(can && live, "null argument");
    KF_CHECK_ARG(same_dims(can->dims, live->dims), "volumes differ in dims");
    KF_CHECK_ARG(can != live, "fusing a volume into itself needs a copy first");
    KF_CHECK_ARG(can->ctx == live->ctx, "handles belong to different contexts");
    KF_CHECK_ARG(mode == KF_FUSE_MERGE || mode == KF_FUSE_NEAREST, "unknown fuse mode");
    DeviceGuard dg(can->ctx->device);
    return kf_launch_fuse(can, live, mode == KF_FUSE_MERGE);
}

kf_status kf_fuse(kf_volume *can, const kf_volume *live) { return kf_fuse_mode(can, live, KF_FUSE_MERGE); }

kf_status kf_fuse_warped_mode(kf_volume *can, const kf_volume *live, const kf_field *u, int32_t mode)
{
    KF_CHECK_ARG(can && live && u, "null argument");
    KF_CHECK_ARG(same_dims(can->dims, live->dims) && same_dims(can->dims, u->dims), "dims differ");
    KF_CHECK_ARG(can != live, "fusing a volume into itself needs a copy first");
    KF_CHECK_ARG(can->ctx == live->ctx && can->ctx == u->ctx, "handles belong to different contexts");
    KF_CHECK_ARG(mode == KF_FUSE_MERGE || mode == KF_FUSE_NEAREST, "unknown fuse mode");
    DeviceGuard dg(can->ctx->device);
    return kf_launch_fuse_warped(can, live, u, mode == KF_FUSE_MERGE);
}

kf_status kf_fuse_warped(kf_volume *can, const kf_volume *live, const kf_field *u)
{
    return kf_fuse_warped_mode(can, live, u, KF_FUSE_MERGE);
}

kf_status kf_marching_cubes(const kf_volume *v, double iso, uint64_t *n_tris, float *verts, uint64_t cap_tris)
{
    KF_CHECK_ARG(v && n_tris, "null argument");
    DeviceGuard dg(v->ctx->device);
    return kf_mc_run(v, iso, n_tris, verts, cap_tris);
}

kf_status kf_marching_cubes_indexed(const kf_volume *v, double iso, uint64_t *n_verts, uint64_t *n_tris, float *verts,
                                    float *normals, uint32_t *tris, uint64_t cap_verts, uint64_t cap_tris)
{
    KF_CHECK_ARG(v && n_verts && n_tris, "null argument");
    KF_CHECK_ARG(3 * ((int64_t)v->dims[0] + 1) * ((int64_t)v->dims[1] + 1) * ((int64_t)v->dims[2] + 1) <
                     ((int64_t)1 << 32),
                 "indexed marching cubes numbers grid edges with 32 bits");
    DeviceGuard dg(v->ctx->device);
    return kf_mc_run_indexed(v, iso, n_verts, n_tris, verts, normals, tris, cap_verts, cap_tris);
}

}  // extern "C"
