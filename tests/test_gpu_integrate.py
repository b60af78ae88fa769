"""Integration (SDF::integrateDepthFrame, SDF.cpp:137-197) under camera poses other than the
identity, against the FP64 oracle bit for bit (weights and the stored rounding of phi).

The kernel decides each voxel's pixel without FP64 divisions wherever a proven margin allows and
takes the reference's FP64 projection elsewhere; the poses reach each of its routes:
  * a rotation about y with a translation: the general form (depth varies along a grid row);
  * a rotation about x: the row form (depth and image row fixed along a grid row);
  * the camera inside the grid, with and without a rotation about y: voxels behind the camera
    and next to its plane, which the exact projection decides (|c2| below the margin's floor),
    and voxels behind it that still project into the image, which the reference integrates.
Each pose runs as an update (SDF.integrateDepthFrame into a reset volume) and as the frame
driver's fresh pass (frame 0 of kf_pipeline: reset and integrate in one write), in FP32, BF16
and FP16 storage.  Grids whose voxel count is not a multiple of four check the per-voxel tails
of the reset, both integration passes and both fusion rules."""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import KillingFusion, KillingParams, SDF, scenes

from helpers import C1, bf16_roundtrip, f16_roundtrip, ogrid, sphere_frames

pytestmark = pytest.mark.gpu

SCALE = np.float32(C1["trunc"] + 1e-5)


def rot(axis, deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    m = np.eye(4)
    i, j = {"x": (1, 2), "y": (0, 2)}[axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def pose(axis, deg, t):
    m = rot(axis, deg)
    m[:3, 3] = t
    return m


POSES = {
    "y-rotation": pose("y", 7.0, (0.02, -0.01, 0.03)),
    "x-rotation": pose("x", -6.0, (0.01, 0.02, -0.02)),
    "inside": pose("y", 0.0, (0.0, 0.0, 0.75)),
    "inside, y-rotation": pose("y", 20.0, (0.05, 0.0, 0.75)),
}


def stored(dtype, phi64):
    if dtype == "f32":
        return np.asarray(phi64, dtype=np.float64).astype(np.float32)
    rt = bf16_roundtrip if dtype == "bf16" else f16_roundtrip
    return rt(np.asarray(phi64, dtype=np.float64).astype(np.float32), SCALE)


def oracle_frame(g, depth, c2w):
    ophi, ow = O.sdf_init(g)
    O.integrate(ophi, ow, g, C1["trunc"], depth, scenes.K, c2w, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    return ophi, ow


def check(phi, w, ophi, ow, dtype, tag):
    assert (w == ow).all(), f"{tag}: {(w != ow).sum()} weight mismatches"
    want = stored(dtype, ophi)
    bad = phi != want
    assert not bad.any(), f"{tag}: {bad.sum()} phi mismatches, e.g. {phi[bad][:3]} vs {want[bad][:3]}"


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("name", list(POSES))
def test_integrate_posed_bit_exact(ctx, name, dtype):
    print(f"case: integrate 64^3 sphere frame, pose {name}, {dtype} storage, update and fresh", flush=True)
    c2w = POSES[name]
    s = sphere_frames()[0]
    depth = np.where(s > 0, s, np.float32(0.5))  # the sphere, and a plane at 0.5 m elsewhere: every pixel valid
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    ophi, ow = oracle_frame(g, depth, c2w)
    assert (ow > 0).sum() > 5000
    # update into a reset volume
    v = SDF(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], C1["trunc"], dtype=dtype)
    v.integrateDepthFrame(depth, c2w, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    phi, w = v.download()
    check(phi, w, ophi, ow, dtype, f"update, {name}")
    v.close()
    # the frame driver's first frame: the fresh pass into the canonical volume
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=1)
    pl = KillingFusion(ctx, C1["dims"], C1["min3d"], C1["voxel"], kp, scenes.K, C2W=c2w,
                       min_depth=scenes.MIN_DEPTH, max_depth=scenes.MAX_DEPTH, dtype=dtype)
    pl.processNextFrame(depth)
    phi, w = pl.canonical_download()
    check(phi, w, ophi, ow, dtype, f"fresh, {name}")
    pl.close()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("dims", [(61, 45, 58), (63, 47, 29), (13, 7, 5)])
def test_uneven_grids_fill_integrate_fuse(ctx, dims, dtype):
    """Grids whose voxel count leaves 2, 1 and 3 voxels past the last group of four: the reset
    (SDF.cpp:133-134), the update and fresh integration, and both fusion rules (SDF.cpp:199-241)
    take their per-voxel tails there."""
    n = int(np.prod(dims))
    print(f"case: uneven grid {dims} (n % 4 = {n % 4}), {dtype}: reset, integrate, fuse", flush=True)
    mn = (-0.5 * dims[0] * C1["voxel"], -0.5 * dims[1] * C1["voxel"], 0.56)  # centred on the sphere
    g = ogrid(dims, mn, C1["voxel"], C1["trunc"])
    d1, d2 = sphere_frames()
    args = (C1["voxel"], C1["trunc"], C1["trunc"])
    v = SDF(ctx, dims, mn, *args, dtype=dtype)
    phi, w = v.download()
    ophi, ow = O.sdf_init(g)
    check(phi, w, ophi, ow, dtype, "reset")
    v.integrateDepthFrame(d1, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    o1, w1 = oracle_frame(g, d1, scenes.C2W)
    phi, w = v.download()
    check(phi, w, o1, w1, dtype, "update")
    assert (w1 > 0).sum() > min(1000, n // 2)
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=1)
    pl = KillingFusion(ctx, dims, mn, C1["voxel"], kp, scenes.K, min_depth=scenes.MIN_DEPTH,
                       max_depth=scenes.MAX_DEPTH, dtype=dtype)
    pl.processNextFrame(d1)
    phi, w = pl.canonical_download()
    check(phi, w, o1, w1, dtype, "fresh")
    pl.close()
    live = SDF(ctx, dims, mn, *args, dtype=dtype)
    live.integrateDepthFrame(d2, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    lphi, lw = live.download()
    for merge in (True, False):
        can = SDF(ctx, dims, mn, *args, dtype=dtype)
        can.copy_from(v)
        can.fuse(live, merge=merge)
        cphi, cw = v.download()
        ophi, ow = cphi.astype(np.float64), cw.astype(np.int64)
        O.fuse(ophi, ow, lphi.astype(np.float64), lw.astype(np.int64), merge=merge)
        phi, w = can.download()
        check(phi, w, ophi, ow, dtype, f"fuse merge={merge}")
        can.close()
    live.close()
    v.close()
