// k_mc_noalloc.hip -- kf_marching_cubes_noalloc: the marching-cubes soup of kf_marching_cubes
// for callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that
// forbid allocation).
//
// The count, scan and emit kernels of k_mc.hip (kf_mc_launch_into) run with their unit and
// block totals and block offsets in `scratch`, a file-scope device array that every GPU holds until the
// process ends, and write the soup and its triangle count to caller device memory.  Nothing is
// allocated and nothing waits on the host.  Host code never passes `scratch` itself to a
// kernel (in host code the name stands for a host-side stand-in, not the device copy): each
// device's address comes from hipGetSymbolAddress, taken once per device when a context is
// created there (kf_noalloc_init), and is what the kernels receive.
//
// The count kernel writes every unit and block total and the scan every offset of the grid it
// runs on, and the list of blocks with triangles with its length, so no part of the array has
// to start at zero (2 MiB: totals, offsets and list for KF_MC_NOALLOC_MAX_VOXELS voxels).
// Every call on one GPU shares its array: calls that overlap on one GPU (other streams, other
// threads) race and must be ordered by the caller.
#include "kf_internal.hpp"

#include <mutex>
#include <string>

namespace {

constexpr int64_t MAX_BLOCKS = (KF_MC_NOALLOC_MAX_VOXELS + KF_MC_CELLS - 1) / KF_MC_CELLS;
constexpr size_t UNIT_WORDS = (size_t)MAX_BLOCKS * (KF_MC_CELLS / KF_MC_UNIT) / 2;  // u32 unit totals
constexpr size_t TOT_WORDS = (size_t)(MAX_BLOCKS + 1) / 2;   // then u32 block totals, two per word
constexpr size_t OFFS_WORDS = (size_t)MAX_BLOCKS + 1;        // then u64 offsets and the total
constexpr size_t LIST_WORDS = (size_t)(MAX_BLOCKS + 2) / 2;  // then the u32 list of blocks with triangles
constexpr size_t SCRATCH_WORDS = UNIT_WORDS + TOT_WORDS + OFFS_WORDS + LIST_WORDS;
static_assert(SCRATCH_WORDS * 8 <= ((size_t)2 << 20) + 256, "scratch stays at 2 MiB");

static __device__ __align__(256) unsigned long long scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
unsigned long long *g_scratch_ptr[MAX_DEVICES];  // scratch's address on each device

}  // namespace

// Records scratch's address on the current device (once per device); kf_context_create calls
// it with the context's device current.
kf_status kf_noalloc_init()
{
    int dev = -1;
    KF_HIP(hipGetDevice(&dev));
    KF_CHECK_ARG(dev >= 0 && dev < MAX_DEVICES, "device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        KF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<unsigned long long *>(ptr);
    }
    return KF_OK;
}

extern "C" kf_status kf_marching_cubes_noalloc(const kf_volume *v, double iso, float *verts, uint64_t cap_tris,
                                               uint64_t *count)
{
    KF_CHECK_ARG(v && count, "null argument");
    KF_CHECK_ARG(verts || cap_tris == 0, "verts is null with a nonzero capacity");
    KF_CHECK_ARG(v->n <= KF_MC_NOALLOC_MAX_VOXELS,
                 "kf_marching_cubes_noalloc: grid above KF_MC_NOALLOC_MAX_VOXELS voxels (its scratch's limit)");
    kf_context *ctx = v->ctx;
    unsigned long long *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (ctx->device >= 0 && ctx->device < MAX_DEVICES)
            base = g_scratch_ptr[ctx->device];
    }
    KF_CHECK_ARG(base, "kf_marching_cubes_noalloc: no scratch address for this device");
    KfDeviceGuard dg(ctx->device);
    unsigned long long *const tot = base + UNIT_WORDS, *const offs = tot + TOT_WORDS, *const list = offs + OFFS_WORDS;
    return kf_mc_launch_into(v, iso, reinterpret_cast<unsigned int *>(base), reinterpret_cast<unsigned int *>(tot),
                             offs, reinterpret_cast<unsigned int *>(list), reinterpret_cast<unsigned long long *>(count),
                             verts, cap_tris);
}
