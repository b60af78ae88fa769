// kf_internal.hpp -- shared host structs, storage codecs and device helpers of libkf.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>

#include "kf_api.h"

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------

void kf_set_error(const std::string &msg);

#define KF_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            kf_set_error(std::string(#call " failed: ") + hipGetErrorString(e_));             \
            return e_ == hipErrorOutOfMemory ? KF_E_OOM : KF_E_HIP;                           \
        }                                                                                     \
    } while (0)

#define KF_CHECK_ARG(cond, msg)                                                               \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            kf_set_error(msg);                                                                \
            return KF_E_INVALID_ARG;                                                          \
        }                                                                                     \
    } while (0)

#define KF_LAUNCH_CHECK()                                                                     \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) {                                                               \
            kf_set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_));      \
            return KF_E_HIP;                                                                  \
        }                                                                                     \
    } while (0)

// ---------------------------------------------------------------------------
// Device-side optimiser state (one per context, reused by every deform call)
// ---------------------------------------------------------------------------

struct DeformState {
    unsigned int done;          // set when the stop rule fires or a voxel diverged
    unsigned int iters;         // iterations completed (launches folded while not done)
    unsigned int run[2];        // adjoint step launch k runs iff run[k & 1]; launch k arms run[(k + 1) & 1]
    unsigned int listed;        // voxels in the active-set list of this deform call
    long long diverged;             // lowest index with a non-finite update, INT64_MAX if none
    long long div_slot[2];          // step launch k takes atomicMin into div_slot[k & 1]; the fold of
                                    // launch k moves it to `diverged` and re-arms it
    double stop;                    // stop threshold (<= 0: never)
    double last_max;                // max |du| of the last completed launch
    double energy;                  // energy of the last completed launch
    unsigned long long last_active; // active voxels of the last completed launch
    unsigned int nonfinite;         // the active-set listing met a non-finite live SDF value
    unsigned int list_n;            // voxels the list kernels step: `listed`, or 0 when list_bad
    double list_energy;             // energy of the final field, active-set sweep
    unsigned int list_bad;          // the list outgrew the capacity its launches were sized for, or
                                    // (reference method) met a non-finite live value: nothing ran
    unsigned int pad2;
};

// Per-block partial results of one step launch, folded by k_finalize.
struct BlockPartial {
    double sum;               // energy
    double max;               // max |du|
    unsigned long long cnt;   // active voxels
    unsigned long long pad;
};

// ---------------------------------------------------------------------------
// Host handles
// ---------------------------------------------------------------------------

// A launch sequence captured as a hipGraph and replayed while its key (everything the captured
// launches read by value: pointers, counts, constants) stays the same.  A sequence runs
// directly the first time a key is seen and is captured when the same key comes back.
struct kf_loop_graph {
    std::string key;   // key of the captured graph
    std::string seen;  // key of the last sequence run directly
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
};

// Device workspace of the active-set sweep, grown on demand and kept by the context.
struct kf_active_ws {
    unsigned long long *mask = nullptr;  // one bit per voxel: listed
    size_t mask_cap = 0;                 // words
    unsigned int *span_count = nullptr;  // listed voxels per span of the marking pass, then span offsets
    size_t span_cap = 0;
    unsigned int *pos_of = nullptr;      // list slot of each listed voxel (other entries unused)
    size_t pos_cap = 0;
    unsigned int *list = nullptr;        // listed voxel indices, ascending
    float4 *u = nullptr;                 // two compact sets of u, [2][cap] (x, y, z, 0)
    int4 *nbr = nullptr;                 // [cap][3]: list slots of the 12 Killing-term neighbours, -1 = +0
    size_t cap = 0;                      // list slots
    // batched sweep: per-volume states and energies, work block -> volume, volume table
    DeformState *bstates = nullptr;
    double *benergy = nullptr;
    DeformState *bstates_host = nullptr;  // pinned: states, then the energies
    size_t bcap = 0;                      // volumes
    unsigned int *blk_vol = nullptr;
    size_t blk_cap = 0;
    unsigned char *vtab = nullptr;
    size_t vtab_cap = 0;                  // bytes
    // batched listing: volume table, and the volume of each block of its launches
    unsigned char *ltab = nullptr;
    size_t ltab_cap = 0;                  // bytes
    unsigned int *span_vol = nullptr;
    size_t span_vol_cap = 0;
    unsigned int *nb_vol = nullptr;
    size_t nb_vol_cap = 0;
    // listing of the voxels a warp changes: counters (device / pinned host) and the new
    // phi, weight of each listed voxel
    unsigned int *aux = nullptr;
    unsigned int *aux_host = nullptr;
    unsigned char *wtmp = nullptr;
    size_t wtmp_cap = 0;                  // bytes
    // The list a kf_deform active-set sweep left behind covers every voxel a warp of `warp_vol`
    // by `warp_field` can change (when no live SDF value was non-finite); the per-frame driver
    // warps over it.  Any other use of the workspace clears it.
    const kf_volume *warp_vol = nullptr;
    const kf_field *warp_field = nullptr;
    unsigned int warp_count = 0;
    // The list length of the last kf_deform call on a volume of hint_n voxels: the next call
    // on such a volume sizes its list launches from it without reading the length back first.
    int64_t hint_n = 0;
    unsigned int hint_listed = 0;
    // launch shape (slots, expected length) of the last predicted call, on a volume of shape_n
    int64_t shape_n = 0;
    unsigned int shape_slots = 0, shape_expect = 0;
    // the iteration loops of kf_deform and kf_batch_deform over the lists (no stop rule)
    kf_loop_graph loop, batch_loop;
};

// Device workspace of the reference-formula step (k_deform_ref.hip), grown on demand and kept
// by the context: the voxels in the band of a step launch, listed by its first pass.
struct kf_ref_ws {
    unsigned long long *idx = nullptr;  // voxel indices, one segment per XCD, in no particular order
    float *psi = nullptr;               // their warped live SDF value
    size_t cap = 0;                     // entries of idx and psi
    unsigned long long *aux = nullptr;  // segment lengths of launch parity 0 and 1, then the non-finite flag
    unsigned int *row_fail = nullptr;   // trust region: first failing x of each (y, z) row, per parity
    size_t rows_cap = 0;                // entries of row_fail
};

// Marching-cubes block offsets of the last volume counted (k_mc.hip): a size query followed by
// the extraction of the same, unchanged volume counts once.
struct kf_mc_cache {
    const kf_volume *vol = nullptr;
    uint64_t version = 0;
    double iso = 0.0;
    bool valid = false;
    unsigned int *unit_tot = nullptr;      // triangles per unit of cells
    unsigned int *tot = nullptr;           // triangles per block
    unsigned long long *offs = nullptr;    // exclusive block offsets, then the total
    unsigned int *blist = nullptr;         // blocks with triangles: their number, then their indices
    size_t cap = 0;                        // blocks
    unsigned long long *total_host = nullptr;  // pinned: the cached triangle total, then a spare slot
    float *verts = nullptr;                // device staging of a soup copied to the host
    size_t verts_cap = 0;                  // triangles
};

struct kf_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    DeformState *state = nullptr;     // device
    DeformState *state_host = nullptr;  // pinned mirror for stats readback
    BlockPartial *partials = nullptr;  // device, two sets (launch parity) of one slot per block
    size_t partials_cap = 0;           // slots per set
    unsigned long long *need_words = nullptr;  // device: the dense step's need bits of its last launch
    size_t need_cap = 0;                       // words
    int num_cus = 256;
    bool capture_loops = false;              // kf_context_set_loop_capture
    struct kf_batch_cache *batch = nullptr;  // batched-deform workspace, tuned shapes, captured graph
    kf_active_ws active;                     // active-set sweep workspace (k_deform_active.hip)
    kf_ref_ws ref;                           // reference-formula step workspace (k_deform_ref.hip)
    kf_mc_cache mc;                          // marching-cubes counts (k_mc.hip)
};

struct kf_volume {
    kf_context *ctx = nullptr;
    int32_t dims[3] = {0, 0, 0};
    int64_t n = 0;
    double min_corner[3] = {0, 0, 0};
    double voxel = 0, clip = 0, max_surface = 0, epsilon = 0;
    kf_dtype dtype = KF_F32;
    void *phi = nullptr;     // device: n elements of dtype
    int32_t *w = nullptr;    // device: n
    // Changes whenever work that writes phi is enqueued (kf_touch); unique across volumes, so
    // results cached for one version (marching-cubes counts) are never taken for another.
    uint64_t version = 0;
    double oob() const { return max_surface + epsilon; }
    // Decode scale: stored value * scale = metres (1 for F32).
    float scale() const { return dtype == KF_F32 ? 1.0f : (float)(max_surface + epsilon); }
    size_t elem_size() const { return dtype == KF_F32 ? 4 : 2; }
};

struct kf_field {
    kf_context *ctx = nullptr;
    int32_t dims[3] = {0, 0, 0};
    int64_t n = 0;
    float *buf[2] = {nullptr, nullptr};  // SoA [3][n] each, one allocation (buf[1] = buf[0] + 3n); buf[cur] is the value
    int cur = 0;
    unsigned int *masks[2] = {nullptr, nullptr};  // stencil + nonzero masks of buf[0] / buf[1] (adjoint kernels)
    float *u() const { return buf[cur]; }
    float *scratch() const { return buf[cur ^ 1]; }
};

// ---------------------------------------------------------------------------
// Storage codecs.  F32 holds metres; BF16/F16 hold phi / (Max + eps).
// ---------------------------------------------------------------------------

// v / scale rounded to float, as (float)((double)v * rscale) where rscale is 1 / scale rounded
// to double (kf_quot_rscale), without a float division.  The two agree for every float v when
// scale is a normal float below 2: the product lies within 2^-52 relative of v / scale, while a
// quotient of two floats that is not a midpoint between adjacent floats lies at least 2^-49
// relative from every midpoint (subnormal results included), and with scale < 2 no quotient is
// a midpoint (its odd significand would need bits below 2^-149).  So both round to the same
// float; inf, NaN and signed zeros pass through the product unchanged.  rscale = 0 (a scale
// outside that range) takes the division.  tests/cpp/integrate_check.cpp compares the two.
__device__ __host__ inline float kf_quot(float v, float scale, double rscale)
{
    return rscale != 0.0 ? (float)((double)v * rscale) : v / scale;
}

inline double kf_quot_rscale(float scale)
{
    return (std::isnormal(scale) && scale > 0.0f && scale < 2.0f) ? 1.0 / (double)scale : 0.0;
}

struct StoreF32 {
    using T = float;
    static constexpr int dtype = KF_F32;
    __device__ __host__ static inline float load(const T *p, int64_t i, float) { return p[i]; }
    // Four consecutive values starting at a multiple of 4 (one 16-byte load).
    __device__ static inline void load4(const T *p, int64_t i, float, float o[4])
    {
        const float4 v = *reinterpret_cast<const float4 *>(p + i);
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
    __device__ __host__ static inline T enc(float v, float) { return v; }
    __device__ __host__ static inline T enc_r(float v, float, double) { return v; }
    __device__ __host__ static inline float dec(T v, float) { return v; }
};

struct StoreBF16 {
    using T = uint16_t;
    static constexpr int dtype = KF_BF16;
    __device__ __host__ static inline float dec(T v, float scale)
    {
        union { uint32_t u; float f; } c;
        c.u = (uint32_t)v << 16;
        return c.f * scale;
    }
    __device__ __host__ static inline float load(const T *p, int64_t i, float scale) { return dec(p[i], scale); }
    // Four consecutive values starting at a multiple of 4 (one 8-byte load).
    __device__ static inline void load4(const T *p, int64_t i, float scale, float o[4])
    {
        const uint2 v = *reinterpret_cast<const uint2 *>(p + i);
        o[0] = __uint_as_float(v.x << 16) * scale;
        o[1] = __uint_as_float(v.x & 0xffff0000u) * scale;
        o[2] = __uint_as_float(v.y << 16) * scale;
        o[3] = __uint_as_float(v.y & 0xffff0000u) * scale;
    }
    // Round to nearest even (NaN kept quiet).
    __device__ __host__ static inline T round(float q)
    {
        union { uint32_t u; float f; } c;
        c.f = q;
        if ((c.u & 0x7fffffffu) > 0x7f800000u)
            return (T)((c.u >> 16) | 0x40u);
        uint32_t r = c.u + 0x7fffu + ((c.u >> 16) & 1u);
        return (T)(r >> 16);
    }
    __device__ __host__ static inline T enc(float v, float scale) { return round(v / scale); }
    // enc with the quotient formed as kf_quot (rscale = 1 / scale rounded to double, or 0).
    __device__ __host__ static inline T enc_r(float v, float scale, double rscale)
    {
        return round(kf_quot(v, scale, rscale));
    }
};

struct StoreF16 {
    using T = _Float16;
    static constexpr int dtype = KF_F16;
    __device__ __host__ static inline float dec(T v, float scale) { return (float)v * scale; }
    __device__ __host__ static inline float load(const T *p, int64_t i, float scale) { return dec(p[i], scale); }
    // Four consecutive values starting at a multiple of 4 (one 8-byte load).
    __device__ static inline void load4(const T *p, int64_t i, float scale, float o[4])
    {
        const uint2 v = *reinterpret_cast<const uint2 *>(p + i);
        o[0] = (float)__builtin_bit_cast(_Float16, (unsigned short)(v.x & 0xffffu)) * scale;
        o[1] = (float)__builtin_bit_cast(_Float16, (unsigned short)(v.x >> 16)) * scale;
        o[2] = (float)__builtin_bit_cast(_Float16, (unsigned short)(v.y & 0xffffu)) * scale;
        o[3] = (float)__builtin_bit_cast(_Float16, (unsigned short)(v.y >> 16)) * scale;
    }
    __device__ __host__ static inline T enc(float v, float scale) { return (T)(v / scale); }
    __device__ __host__ static inline T enc_r(float v, float scale, double rscale)
    {
        return (T)kf_quot(v, scale, rscale);
    }
};

// Four consecutive stored SDF values from p, decoded, where p is aligned to its element size
// only: one 8-byte (BF16/F16) or 16-byte (F32) load.
template <class S>
__device__ inline void kf_load4_unaligned(const typename S::T *p, float scale, float o[4])
{
    typedef unsigned int raw_t
        __attribute__((ext_vector_type(sizeof(typename S::T)), aligned(sizeof(typename S::T))));
    static_assert(sizeof(raw_t) == 4 * sizeof(typename S::T), "four elements per raw load");
    typename S::T t[4] __attribute__((aligned(16)));
    const raw_t v = *reinterpret_cast<const raw_t *>(p);
    __builtin_memcpy(t, &v, sizeof(t));
    S::load4(t, 0, scale, o);
}

// Dispatch a templated functor on a volume's dtype.
#define KF_DISPATCH_DTYPE(dt, S, ...)                                                         \
    switch (dt) {                                                                             \
    case KF_F32: { using S = StoreF32; __VA_ARGS__; } break;                                  \
    case KF_BF16: { using S = StoreBF16; __VA_ARGS__; } break;                                \
    case KF_F16: { using S = StoreF16; __VA_ARGS__; } break;                                  \
    default: kf_set_error("unknown dtype"); return KF_E_INVALID_ARG;                          \
    }

// ---------------------------------------------------------------------------
// Grid geometry passed to kernels by value
// ---------------------------------------------------------------------------

// Unsigned 32-bit division by a divisor fixed before a launch, as a multiply-high and two
// shifts (Granlund and Montgomery, "Division by Invariant Integers using Multiplication",
// PLDI 1994, figure 4.1): with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// s1 = min(l, 1), s2 = max(l - 1, 0) and t = the high word of m n,
// n / d = (t + ((n - t) >> s1)) >> s2 exactly, for every n < 2^32.  Five integer instructions
// where a division by a value known only at run time takes about twenty.
struct U32Div {
    uint32_t m = 1, s1 = 0, s2 = 0;  // d = 1
    static inline U32Div of(uint32_t d)  // d >= 1
    {
        U32Div r;
        uint32_t l = 0;
        while (l < 32 && ((uint64_t)1 << l) < d)
            ++l;
        // 2^l - d < d, so m < 2^32
        r.m = (uint32_t)(((((uint64_t)1 << l) - d) << 32) / d + 1);
        r.s1 = l < 1 ? l : 1;
        r.s2 = l > 0 ? l - 1 : 0;
        return r;
    }
    __device__ __host__ inline uint32_t div(uint32_t n) const
    {
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

struct GridDims {
    int nx, ny, nz;
    int64_t n;
    U32Div dx, dxy;  // division by nx and by nx ny (the latter only when nx ny < 2^32)
    __device__ __host__ inline bool in(int x, int y, int z) const
    {
        return x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
    }
    __device__ __host__ inline int64_t idx(int x, int y, int z) const
    {
        return ((int64_t)z * ny + y) * nx + x;
    }
    // 32-bit index for kernels restricted to fields of fewer than 2^30 floats per buffer.
    __device__ __host__ inline unsigned int idx32(int x, int y, int z) const
    {
        return ((unsigned int)z * (unsigned int)ny + (unsigned int)y) * (unsigned int)nx + (unsigned int)x;
    }
    // (x, y, z) of voxel index i of a grid of fewer than 2^32 voxels.
    __device__ __host__ inline void xyz32(uint32_t i, int &x, int &y, int &z) const
    {
        const uint32_t zz = dxy.div(i), r = i - zz * ((uint32_t)nx * (uint32_t)ny), yy = dx.div(r);
        z = (int)zz;
        y = (int)yy;
        x = (int)(r - yy * (uint32_t)nx);
    }
};

inline GridDims make_grid(int nx, int ny, int nz, int64_t n)
{
    GridDims g{nx, ny, nz, n};
    g.dx = U32Div::of((uint32_t)nx);
    const uint64_t nxy = (uint64_t)nx * (uint64_t)ny;
    g.dxy = nxy < ((uint64_t)1 << 32) ? U32Div::of((uint32_t)nxy) : U32Div{};
    return g;
}

// A fresh version for a volume whose phi is about to change (see kf_volume::version).
uint64_t kf_next_version();
inline void kf_touch(kf_volume *v) { v->version = kf_next_version(); }

inline GridDims grid_of(const kf_volume *v) { return make_grid(v->dims[0], v->dims[1], v->dims[2], v->n); }
inline GridDims grid_of(const kf_field *f) { return make_grid(f->dims[0], f->dims[1], f->dims[2], f->n); }

inline bool same_dims(const int32_t a[3], const int32_t b[3]) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }

// Makes a context's device current for the scope of an entry point and restores the caller's.
struct KfDeviceGuard {
    int prev = -1;
    explicit KfDeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        if (prev != dev)
            (void)hipSetDevice(dev);
    }
    ~KfDeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev)
            (void)hipSetDevice(prev);
    }
    KfDeviceGuard(const KfDeviceGuard &) = delete;
    KfDeviceGuard &operator=(const KfDeviceGuard &) = delete;
};

// (x, y, z) of linear voxel index i: multiply-high divisions (U32Div) when N < 2^32, 64-bit
// integer divisions (a long emulated sequence) otherwise.
__device__ inline void kf_xyz(const GridDims &g, int64_t i, int &x, int &y, int &z)
{
    if (g.n < ((int64_t)1 << 32)) {
        g.xyz32((uint32_t)i, x, y, z);
    } else {
        x = (int)(i % g.nx);
        y = (int)((i / g.nx) % g.ny);
        z = (int)(i / ((int64_t)g.nx * g.ny));
    }
}

// MI355X dispatches workgroup b to XCD (b mod 8), and each XCD has its own L2.  This maps
// the dispatch index to a tile index so that every XCD sweeps one contiguous range of tiles:
// tiles that share halo rows and slices then run on the same XCD and hit its L2.
constexpr int KF_NUM_XCD = 8;
__device__ __host__ inline unsigned int kf_xcd_tile(unsigned int b, unsigned int nb)
{
    const unsigned int xcd = b % KF_NUM_XCD, local = b / KF_NUM_XCD;
    const unsigned int per = nb / KF_NUM_XCD, rem = nb % KF_NUM_XCD;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Chunks of C consecutive tiles dealt round-robin over the XCDs (XCD x sweeps chunks x, x + 8,
// ... in order), the tiles left over by whole rounds of 8 C as kf_xcd_tile's contiguous
// ranges.  Within a chunk neighbouring tiles share an L2; across the grid every XCD gets an
// equal share of any spatially concentrated work, such as the stencils of a surface band that
// a contiguous split would hand to two or three XCDs.
template <unsigned int C>
__device__ __host__ inline unsigned int kf_xcd_chunk_tile(unsigned int b, unsigned int nb)
{
    const unsigned int xcd = b % KF_NUM_XCD, local = b / KF_NUM_XCD;
    const unsigned int K = nb / (KF_NUM_XCD * C);
    if (local < K * C)
        return (local / C) * (KF_NUM_XCD * C) + xcd * C + local % C;
    const unsigned int rl = local - K * C, r = nb - K * KF_NUM_XCD * C;
    const unsigned int per = r / KF_NUM_XCD, rem = r % KF_NUM_XCD;
    return K * KF_NUM_XCD * C + xcd * per + (xcd < rem ? xcd : rem) + rl;
}

// Base index and weight of the sample at i + r (i an integer grid index, r a small offset),
// with the reference's truncating cast: base = (int)(i + r), frac = (i + r) - base
// (SDF.cpp:509, DisplacementField.cpp:43).  i + r is never formed in FP32: at i ~ 200 that
// would round r to ~1e-5 voxel and move samples across cell faces; taking the floor of r
// alone keeps the weight accurate to the precision of r.
__device__ __host__ inline void kf_split(int i, float r, int &b, float &f)
{
    const float fl = floorf(r);
    // The cell index is clamped to +-2^30 before the conversion (the weight keeps the exact
    // floor): for |r| >= 2^30 every tap lies outside the grid (at most 2^15 per axis) and no
    // index arithmetic can overflow.  The reference's own (int) cast is undefined there.
    int bi = i + (int)fminf(fmaxf(fl, -1073741824.0f), 1073741824.0f);
    float fr = r - fl;
    if (bi < 0 && fr > 0.0f) {  // i + r < 0, not an integer: truncation rounds toward zero
        bi += 1;
        fr -= 1.0f;
    }
    b = bi;
    f = fr;
}

// Interpolation chain of utils.h:6-46 (x, then y, then z; v0*(1-x) + v1*x).
template <typename R>
__device__ __host__ inline R kf_lerp(R v0, R v1, R x) { return v0 * (R(1) - x) + v1 * x; }

template <typename R>
__device__ __host__ inline R kf_interp3(R v000, R v001, R v010, R v011, R v100, R v101, R v110, R v111,
                                        R x, R y, R z)
{
    R s = kf_lerp(kf_lerp(v000, v001, x), kf_lerp(v010, v011, x), y);
    R t = kf_lerp(kf_lerp(v100, v101, x), kf_lerp(v110, v111, x), y);
    return kf_lerp(s, t, z);
}

// ---------------------------------------------------------------------------
// Launch helpers (defined in the kernel translation units)
// ---------------------------------------------------------------------------

kf_status kf_ensure_partials(kf_context *ctx, size_t nblocks);
kf_status kf_ensure_need_words(kf_context *ctx, size_t words);
bool kf_valid_params(const kf_params *p);
kf_status kf_check_triplet(const kf_volume *live, const kf_volume *can, const kf_field *u);
void kf_batch_release(kf_context *ctx);
inline BlockPartial *kf_partials(kf_context *ctx, int parity) { return ctx->partials + (size_t)(parity & 1) * ctx->partials_cap; }
kf_status kf_state_reset(kf_context *ctx, double stop);
// Folds the nb block partials of the launch with this parity into the iteration state
// (one block, fixed order).
kf_status kf_launch_finalize(kf_context *ctx, unsigned int nb, int parity);

kf_status kf_launch_fill_volume(kf_volume *v);
kf_status kf_launch_upload_volume(kf_volume *v, const float *dphi_f32, const int32_t *dw);
kf_status kf_launch_download_volume(const kf_volume *v, float *dphi_f32);
// fresh: reset the volume (phi = Max + eps, w = 0) in the same pass, before integrating.
kf_status kf_launch_integrate(kf_volume *v, const float *ddepth, int w, int h, const double K[9],
                              const double W2C[16], double min_depth, double max_depth, bool fresh = false);
// kf_integrate_depth, optionally into a volume reset in the same pass.
kf_status kf_integrate_depth_impl(kf_volume *v, const float *depth, int32_t width, int32_t height, const double K[9],
                                  const double C2W[16], double min_depth, double max_depth, bool fresh);
kf_status kf_launch_warp(kf_volume *v, const kf_field *u, void *tmp_phi, int32_t *tmp_w);
// merge: FUSE_BY_MERGE (config.cpp:7), the weighted average; otherwise the value nearer the
// surface wins (SDF.cpp:226-237, 285-295).
kf_status kf_launch_fuse(kf_volume *can, const kf_volume *live, bool merge = true);
kf_status kf_launch_fuse_warped(kf_volume *can, const kf_volume *live, const kf_field *u, bool merge = true);
kf_status kf_launch_field_fill(kf_field *f, const double v[3]);
kf_status kf_launch_field_add(kf_field *dst, const kf_field *src);

struct DeformArgs;
// Step launches take their index `iter` within a deform call: launch k writes partial set k & 1.
// The reference step folds its own partials right after the launch.  vlist: the active-set
// sweep's list of the voxels that can move (step mode only; state->list_n entries, launched
// over vslots slots), whose pass 1 then visits only them; every other voxel must hold +0 in
// both u_in and u_out.
kf_status kf_launch_ref_step(const kf_volume *live, const kf_volume *can, const float *u_in,
                             float *u_out, const kf_params *p, int mode, int iter = 0,
                             const unsigned int *vlist = nullptr, unsigned int vslots = 0);
// mask_in: stencil mask of u_in (kf_launch_adj_mask), needed by energy and gradient passes of
// the row kernel (see k_deform_adj.hip); without it those passes run one voxel per thread.
// An adjoint step launch k > 0 folds launch k - 1 in its first block, so a deform loop ends
// with kf_launch_finalize(ctx, *nb_out, last & 1).  Energy passes fold themselves.
kf_status kf_launch_adj_step(const kf_volume *live, const kf_volume *can, const float *u_in,
                             float *u_out, const kf_params *p, int mode, float *e_voxel,
                             const unsigned int *mask_in = nullptr, int iter = 0,
                             unsigned int *nb_out = nullptr);
// Active-set sweep (k_deform_active.hip), one deform call.  The list's kernels are launched
// over `slots` list slots (a multiple of 1024, kf_list_slots) chosen before its length is known
// on the host; they read the length from state->list_n.  kf_active_count marks the voxels that
// can move and leaves their number in state->listed (reset the state first); when it exceeds
// `slots` (or, with need_finite, a live value is not finite) it sets state->list_bad and the
// list kernels do nothing.  kf_active_prepare lists them and copies u into compact set 0;
// kf_launch_adj_list is step launch `iter` (compact set iter & 1 -> the other), with the same
// launch-to-launch folding as kf_launch_adj_step; kf_active_scatter writes compact set
// iters & 1 back into u.
bool kf_active_supported(const kf_volume *live);
void kf_active_release(kf_context *ctx);
void kf_ref_release(kf_context *ctx);
// Slots for a list of `count` voxels: count rounded up to the quantum, or, with headroom (a
// length predicted from an earlier call), an eighth more rounded to a sixteenth of its power
// of two.  0 when that is more than KF_LIST_MAX (the list kernels index the compact sets'
// floats with 32-bit offsets: 4 per slot).
constexpr unsigned int KF_LIST_MAX = 1u << 30;
unsigned int kf_list_slots(unsigned int count, bool headroom);
kf_status kf_active_count(const kf_volume *live, const float *u, const kf_params *p, bool u_zero, unsigned int slots,
                          bool need_finite);
// neighbours = false lists the voxels only (ctx->active.list), without the adjoint step's
// compact u and neighbour slots.  When the workspace has to grow, it grows to `reserve` slots
// if that is more (so later calls with somewhat longer lists need not grow it again).
kf_status kf_active_prepare(const kf_volume *live, const float *u, unsigned int slots, bool neighbours = true,
                            unsigned int reserve = 0);
// expect (a multiple of 1024, at most slots): the list length the launch's blocks are dealt
// over the XCDs for (kf_list_slots of the predicted or read-back length).
kf_status kf_launch_adj_list(const kf_volume *live, const kf_volume *can, const kf_params *p, int iter,
                             unsigned int slots, unsigned int expect, unsigned int *nb_out);
// All p->max_iters step launches of a call without a stop rule and the fold of the last one,
// replayed as a captured hipGraph when the same launch shape comes back (kf_loop_graph).
kf_status kf_active_loop(const kf_volume *live, const kf_volume *can, const kf_params *p, unsigned int slots,
                         unsigned int expect);
// Runs launches() directly, or, when the context captures loops, as the hipGraph g holds for
// key (capturing it on the key's second run); launches() must only launch kernels on the
// context's stream (no allocation, no memset, no host wait).
template <class F>
kf_status kf_run_captured(kf_context *ctx, kf_loop_graph &g, const std::string &key, F launches)
{
    if (!ctx->capture_loops)
        return launches();
    if (g.exec && key == g.key) {
        KF_HIP(hipGraphLaunch(g.exec, ctx->stream));
        return KF_OK;
    }
    if (key != g.seen) {
        g.seen = key;
        return launches();
    }
    g.drop();
    KF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    kf_status s = launches();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    if (s != KF_OK || ec != hipSuccess || !graph) {
        if (graph)
            (void)hipGraphDestroy(graph);
        if (s == KF_OK) {
            kf_set_error(std::string("loop capture: ") + hipGetErrorString(ec));
            s = KF_E_HIP;
        }
        return s;
    }
    g.graph = graph;
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        g.drop();
        kf_set_error(std::string("loop graph instantiate: ") + hipGetErrorString(ei));
        return KF_E_HIP;
    }
    g.key = key;
    KF_HIP(hipGraphLaunch(g.exec, ctx->stream));
    return KF_OK;
}
// Writes the final compact set (state->iters & 1, read on the device) back into u.
kf_status kf_active_scatter(const kf_volume *live, float *u, unsigned int slots);
// kf_batch_deform over the voxels that can move in each volume (one launch per iteration for
// the whole batch).  Sets *used = false and does nothing more when, without `force`, the
// lists hold more than 1/8 of the voxels.
kf_status kf_batch_active(kf_context *ctx, const kf_volume *const *live, const kf_volume *const *can,
                          kf_field *const *u, int count, const kf_params *p, kf_deform_stats *stats, bool force,
                          bool *used);
// Lists, in ctx->active.list (ascending), the voxels SDF::update can change given u (see
// k_moved_mark); reads back their number and whether some phi is not finite (then nothing
// is listed and the caller must warp every voxel).
kf_status kf_list_moved(const kf_volume *v, const float *u, unsigned int *count, bool *nonfinite);
// SDF::update over those voxels only, in place (k_volume.hip).  Sets *done = false, changing
// nothing, when the list holds more than 1/8 of the voxels or some phi is not finite.
kf_status kf_warp_sparse(kf_volume *v, const kf_field *u, bool *done);
// Energy of the final compact set into state->list_energy (no other state field changes).
kf_status kf_active_energy(const kf_volume *live, const kf_volume *can, const kf_params *p, unsigned int slots,
                           unsigned int expect);
// SDF::update of v by u over the list a kf_deform active-set sweep of (v, u) just left behind
// (k_volume.hip); *done = false, changing nothing, when there is none.
kf_status kf_warp_after_deform(kf_volume *v, const kf_field *u, bool *done);
// kf_deform; u_zero: the caller has just set every u to +0 (the listing skips reading u).
kf_status kf_deform_impl(const kf_volume *live, const kf_volume *can, kf_field *u, const kf_params *p,
                         kf_deform_stats *stats, bool u_zero);
// Words of a nonzero mask for these dims (0: the row kernel does not apply).
size_t kf_adj_mask_words(const int32_t dims[3]);
// Computes the stencil mask of field values u into mask[0, words): every bit whose group's
// adjoint stencil reads a nonzero u.  mask[words, 2 words) holds the nonzero mask it is
// dilated from.
kf_status kf_launch_adj_mask(const kf_field *f, const float *u, unsigned int *mask);

kf_status kf_mc_run(const kf_volume *v, double iso, uint64_t *n_tris, float *verts_host, uint64_t cap);
// The triangle soup of v into device memory: *dverts (capacity *dverts_cap triangles, grown with
// hipMalloc when short), *n_tris the count.  Stream-ordered; one host wait for the count.
kf_status kf_mc_to_device(const kf_volume *v, double iso, float **dverts, size_t *dverts_cap, uint64_t *n_tris);
void kf_mc_release(kf_context *ctx);
// Cells per block of the marching-cubes count pass (one triangle total, one offset and one
// entry of the list of blocks with triangles each), and per unit (the cells of one of its
// wavefronts: one triangle total and one workgroup of the emit pass each).  A grid of n voxels
// has kf_mc_blocks(n) = ceil(n / KF_MC_CELLS) blocks of KF_MC_CELLS / KF_MC_UNIT units.
constexpr int KF_MC_CELLS = 2048;
constexpr int KF_MC_UNIT = 512;
int64_t kf_mc_blocks(int64_t n);
// Count, scan and (cap > 0) emit of v's triangle soup with caller-provided device buffers
// (nb = kf_mc_blocks(n)): unit_tot[4 nb] (16-byte aligned), tot[nb], offs[nb + 1],
// blist[nb + 1], verts[9 cap]; *total_out (may be null, device) receives the triangle count.
// Stream-ordered kernels only: no allocation, no host wait.
kf_status kf_mc_launch_into(const kf_volume *v, double iso, unsigned int *unit_tot, unsigned int *tot,
                            unsigned long long *offs, unsigned int *blist, unsigned long long *total_out,
                            float *verts, uint64_t cap);
kf_status kf_mc_run_indexed(const kf_volume *v, double iso, uint64_t *n_verts, uint64_t *n_tris, float *verts_host,
                            float *normals_host, uint32_t *tris_host, uint64_t cap_verts, uint64_t cap_tris);

// Step-kernel modes.
enum { KF_MODE_STEP = 0, KF_MODE_GRAD = 1, KF_MODE_ENERGY = 2 };
