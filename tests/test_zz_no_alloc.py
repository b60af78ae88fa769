"""kf_marching_cubes_noalloc (the marching cubes entry point that allocates nothing) through the
C ABI, against kf_marching_cubes on the same volumes: the same triangle count and the same soup
bit for bit, on two calls in a row, the second on a smaller grid (the shared scratch keeps the
first call's contents), plus a clipped capacity, a count-only call and the size limit."""
import ctypes as C

import numpy as np
import pytest

from killingfusion_amd import SDF, _lib, scenes
from killingfusion_amd._lib import KF_E_INVALID_ARG, KF_OK

from helpers import C1, make_pair, sphere_frames

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory from the HIP runtime (the caller's allocation, not the library's)."""
    hip = None

    def __init__(self, nbytes):
        if DeviceBuffer.hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            DeviceBuffer.hip = h
        self.ptr = C.c_void_p()
        self.nbytes = max(int(nbytes), 8)
        assert DeviceBuffer.hip.hipMalloc(C.byref(self.ptr), self.nbytes) == 0

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if count:
            assert DeviceBuffer.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def fill(self, byte):
        host = np.full(self.nbytes, byte, dtype=np.uint8)
        assert DeviceBuffer.hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0  # host to device

    def free(self):
        DeviceBuffer.hip.hipFree(self.ptr)


def noalloc(ctx, vol, verts, cap, count):
    L = _lib.load()
    st = L.kf_marching_cubes_noalloc(vol.handle, 0.0, verts.ptr if verts else None, cap, count.ptr)
    ctx.synchronize()
    return st


def test_noalloc_matches_kf_marching_cubes_twice(ctx):
    print("case: kf_marching_cubes_noalloc, 64^3 sphere (f32), then a 40^3 bf16 volume", flush=True)
    d1, d2 = sphere_frames()
    big, _ = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1)
    small = SDF(ctx, (40, 40, 40), (-0.16, -0.16, 0.64), C1["voxel"], C1["trunc"], C1["trunc"], dtype="bf16")
    small.integrateDepthFrame(d1, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    ref_big, ref_small = big.getMesh().tris, small.getMesh().tris
    assert big.n > small.n and min(ref_big.shape[0], ref_small.shape[0]) > 1000
    cap = max(ref_big.shape[0], ref_small.shape[0]) + 16
    verts, count = DeviceBuffer(cap * 36), DeviceBuffer(8)
    for vol, ref in ((big, ref_big), (small, ref_small)):
        verts.fill(0xFF)  # NaN everywhere: every triangle slot of the result must be written
        assert noalloc(ctx, vol, verts, cap, count) == KF_OK
        n = int(count.download(np.uint64, 1)[0])
        got = verts.download(np.float32, 9 * n).reshape(n, 3, 3)
        print(f"{vol.dims}: {n} triangles", flush=True)
        assert n == ref.shape[0]
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # capacity below the count: the count is whole, the first cap triangles are written
    verts.fill(0xFF)
    assert noalloc(ctx, big, verts, 100, count) == KF_OK
    assert int(count.download(np.uint64, 1)[0]) == ref_big.shape[0]
    assert np.array_equal(verts.download(np.float32, 900).reshape(100, 3, 3), ref_big[:100])
    assert np.isnan(verts.download(np.float32, 909)[900:]).all()  # nothing past the capacity
    # count only
    assert noalloc(ctx, small, None, 0, count) == KF_OK
    assert int(count.download(np.uint64, 1)[0]) == ref_small.shape[0]
    verts.free()
    count.free()
    big.close()
    small.close()


def test_noalloc_rejects_grids_above_its_limit(ctx):
    print("case: kf_marching_cubes_noalloc on 512 x 512 x 513 voxels (above 2^27)", flush=True)
    huge = SDF(ctx, (512, 512, 513), (0.0, 0.0, 0.0), 0.01, 0.03, 0.03, dtype="bf16")
    count = DeviceBuffer(8)
    L = _lib.load()
    assert L.kf_marching_cubes_noalloc(huge.handle, 0.0, None, 0, count.ptr) == KF_E_INVALID_ARG
    assert b"KF_MC_NOALLOC_MAX_VOXELS" in L.kf_last_error()
    count.free()
    huge.close()
