"""The per-frame volume kernels on BF16 and FP16 SDF storage against the FP64 oracle.

A BF16/FP16 volume stores round(phi / (Max + eps)) and decodes to FP32 metres; the kernels
compute in FP64 from the decoded values, exactly as on FP32 storage, and round the result
once when they store it.  So each kernel must equal the oracle run on the downloaded
(decoded) grids, with its FP64 result put through the same rounding -- bit for bit, weights
and triangle soups included:
  integrate (SDF.cpp:137-197), on a fresh volume and into one that already holds a frame;
  fuse (SDF.cpp:199-241) and fuse with a displacement field (SDF.cpp:244-298);
  warp (SDF.cpp:300-328), dense and over the moved voxels only;
  marching cubes (SDF.cpp:330-414, MarchingCubes.h:326-422).
"""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import SDF, DisplacementField, scenes

from helpers import C1, bf16_roundtrip, f16_roundtrip, make_pair, ogrid, sphere_frames

pytestmark = pytest.mark.gpu

DTYPES = ["bf16", "f16"]
SCALE = np.float32(C1["trunc"] + 1e-5)


def stored(dtype, phi64):
    """What a volume of this dtype holds after storing the FP64 result phi64 (as the kernels
    do: narrow to FP32, then encode)."""
    rt = bf16_roundtrip if dtype == "bf16" else f16_roundtrip
    return rt(np.asarray(phi64, dtype=np.float64).astype(np.float32), SCALE)


def volume(ctx, dtype):
    return SDF(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], C1["trunc"], dtype=dtype)


@pytest.fixture(scope="module")
def pairs(ctx):
    d1, d2 = sphere_frames()
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    out = {dt: make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype=dt) for dt in DTYPES}
    return dict(g=g, d1=d1, d2=d2, vols=out)


def assert_volume(vol, ophi_stored, ow, tag):
    phi, w = vol.download()
    assert (w == ow).all(), f"{tag}: {(w != ow).sum()} weight mismatches"
    bad = phi != ophi_stored
    assert not bad.any(), f"{tag}: {bad.sum()} phi mismatches, e.g. {phi[bad][:3]} vs {ophi_stored[bad][:3]}"


@pytest.mark.parametrize("dtype", DTYPES)
def test_integrate_bit_exact_low_precision(ctx, pairs, dtype):
    print(f"case: integrate 64^3 sphere frames, {dtype} storage (fresh, then a second frame)", flush=True)
    g = pairs["g"]
    live, can = pairs["vols"][dtype]
    for vol, d in ((live, pairs["d2"]), (can, pairs["d1"])):
        ophi, ow = O.sdf_init(g)
        O.integrate(ophi, ow, g, C1["trunc"], d, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        assert_volume(vol, stored(dtype, ophi), ow, f"integrate {dtype}")
        assert (ow > 0).sum() > 10000
    # a second frame into a volume that holds one: the running average starts from the
    # decoded stored value (SDF.cpp:189-193)
    v = volume(ctx, dtype)
    v.copy_from(live)
    v.integrateDepthFrame(pairs["d1"], scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    phi0, w0 = live.download()
    ophi, ow = phi0.astype(np.float64), w0.astype(np.int64)
    O.integrate(ophi, ow, g, C1["trunc"], pairs["d1"], scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    assert (ow > 1).sum() > 10000
    assert_volume(v, stored(dtype, ophi), ow, f"integrate twice {dtype}")
    v.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_fuse_and_warp_bit_exact_low_precision(ctx, pairs, dtype):
    print(f"case: fuse / fuse_warped / warp 64^3, {dtype} storage", flush=True)
    g = pairs["g"]
    live, can = pairs["vols"][dtype]
    lphi, lw = live.download()
    cphi, cw = can.download()
    u = scenes.seeded_field(C1["dims"], amplitude=1.5, seed=21)
    field = DisplacementField(ctx, C1["dims"])
    field.upload(u)
    a = volume(ctx, dtype)
    # fuse
    a.copy_from(can)
    a.fuse(live)
    ophi, ow = cphi.astype(np.float64), cw.astype(np.int64)
    O.fuse(ophi, ow, lphi.astype(np.float64), lw.astype(np.int64))
    assert_volume(a, stored(dtype, ophi), ow, f"fuse {dtype}")
    # fuse twice (weights above 1 on both sides of the average)
    fphi, fw = a.download()
    a.fuse(live)
    ophi, ow = fphi.astype(np.float64), fw.astype(np.int64)
    O.fuse(ophi, ow, lphi.astype(np.float64), lw.astype(np.int64))
    assert (ow > 2).sum() > 1000
    assert_volume(a, stored(dtype, ophi), ow, f"fuse twice {dtype}")
    # fuse_warped
    a.copy_from(can)
    a.fuse(live, field)
    ophi, ow = cphi.astype(np.float64), cw.astype(np.int64)
    O.fuse_warped(ophi, ow, lphi.astype(np.float64), lw.astype(np.int64), u.astype(np.float64), g)
    assert_volume(a, stored(dtype, ophi), ow, f"fuse_warped {dtype}")
    # warp, dense (u nonzero everywhere)
    a.copy_from(live)
    a.update(field)
    ophi, ow = lphi.astype(np.float64), lw.astype(np.int64)
    O.warp(ophi, ow, u.astype(np.float64), g)
    assert_volume(a, stored(dtype, ophi), ow, f"warp {dtype}")
    # warp over the moved voxels only (u nonzero on a sparse set)
    us = u.copy()
    us[:, scenes.uniform(5, live.n) > 0.01] = 0.0
    field.upload(us)
    a.copy_from(live)
    a.update(field)
    ophi, ow = lphi.astype(np.float64), lw.astype(np.int64)
    O.warp(ophi, ow, us.astype(np.float64), g)
    assert_volume(a, stored(dtype, ophi), ow, f"sparse warp {dtype}")
    a.close()
    field.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_marching_cubes_bit_exact_low_precision(ctx, pairs, dtype):
    print(f"case: marching cubes 64^3 live, canonical and fused, {dtype} storage", flush=True)
    live, can = pairs["vols"][dtype]
    fused = volume(ctx, dtype)
    fused.copy_from(can)
    fused.fuse(live)
    for vol in (live, can, fused):
        phi, _ = vol.download()
        mesh = vol.getMesh()
        ref = O.marching_cubes(phi.astype(np.float64), pairs["g"], C1["trunc"], C1["trunc"])
        assert mesh.tris.shape[0] == ref.shape[0] > 1000
        assert np.array_equal(mesh.tris, ref.astype(np.float32))
    fused.close()


def test_volume_copy_across_contexts_is_rejected(ctx, pairs):
    from killingfusion_amd import Context
    from killingfusion_amd._lib import KF_E_INVALID_ARG, KFError
    print("case: kf_volume_copy between two contexts", flush=True)
    other = Context(0)
    v = SDF(other, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], C1["trunc"], dtype="bf16")
    with pytest.raises(KFError) as e:
        v.copy_from(pairs["vols"]["bf16"][0])
    assert e.value.status == KF_E_INVALID_ARG
    other.close()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("dims", [(61, 45, 58), (63, 2, 3), (5, 9, 7)])
def test_integrate_uneven_rows_and_held_values(ctx, dims, dtype):
    """Rows that are not a multiple of the four voxels a thread integrates (its group wraps into
    the next row), fresh and into held values: weights 0..2, and untouched voxels holding +-inf
    and NaN with weight 0, which the running average turns into NaN (SDF.cpp:189-193)."""
    print(f"case: integrate sphere frame into {dims}, {dtype}, fresh and into held values", flush=True)
    d1, _ = sphere_frames()
    vox = 0.3 / max(dims)
    mn = (-0.15, -0.15, 0.65)
    g = ogrid(dims, mn, vox, 3 * vox)
    rng = np.random.default_rng(sum(dims))
    n = g.n
    scale = np.float32(3 * vox + 1e-5)
    rt = bf16_roundtrip if dtype == "bf16" else (lambda a, s: np.asarray(a, np.float32))
    v = SDF(ctx, dims, mn, vox, 3 * vox, 3 * vox, dtype=dtype)
    v.integrateDepthFrame(d1, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    ophi, ow = O.sdf_init(g)
    O.integrate(ophi, ow, g, 3 * vox, d1, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    phi, w = v.download()
    assert (w == ow).all() and np.array_equal(phi, rt(ophi.astype(np.float32), scale))
    held = rng.uniform(-0.02, 0.02, n).astype(np.float32)
    hw = rng.integers(0, 3, n).astype(np.int32)
    held[:7] = [np.inf, -np.inf, np.nan, np.inf, -np.inf, np.nan, 0.0]
    hw[:7] = 0
    v.upload(held, hw)
    hphi, hw2 = v.download()
    v.integrateDepthFrame(d1, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    ophi, ow = hphi.astype(np.float64), hw2.astype(np.int64)
    O.integrate(ophi, ow, g, 3 * vox, d1, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    phi, w = v.download()
    assert (w == ow).all()
    assert np.array_equal(phi, rt(ophi.astype(np.float32), scale), equal_nan=True)
    v.close()


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_fuse_nearest_mode_bit_exact(ctx, dtype):
    """FUSE_BY_MERGE false (config.cpp:7; SDF.cpp:226-237, 285-295): per voxel the value nearer
    the surface wins with its own weight; plain and displacement-aware fuse against the oracle."""
    print(f"case: fuse / fuse_warped, nearest-surface mode, 64^3, {dtype}", flush=True)
    d1, d2 = sphere_frames()
    live, can = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype=dtype)
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    rt = {"f32": lambda a: np.asarray(a, np.float32), "bf16": lambda a: stored("bf16", a),
          "f16": lambda a: stored("f16", a)}[dtype]
    # a canonical SDF holding two frames (weights 2) and the live one
    can.fuse(live)
    lphi, lw = live.download()
    cphi, cw = can.download()
    u = scenes.seeded_field(C1["dims"], amplitude=1.2, seed=33)
    field = DisplacementField(ctx, C1["dims"])
    field.upload(u)
    a = SDF(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], C1["trunc"], dtype=dtype)
    a.copy_from(can)
    a.fuse(live, merge=False)
    ophi, ow = cphi.astype(np.float64), cw.astype(np.int64)
    O.fuse(ophi, ow, lphi.astype(np.float64), lw.astype(np.int64), merge=False)
    assert (ow != cw).any() and (ow == cw).any()  # both branches taken
    gphi, gw = a.download()
    assert (gw == ow).all() and np.array_equal(gphi, rt(ophi.astype(np.float32)))
    a.copy_from(can)
    a.fuse(live, field, merge=False)
    ophi, ow = cphi.astype(np.float64), cw.astype(np.int64)
    O.fuse_warped(ophi, ow, lphi.astype(np.float64), lw.astype(np.int64), u.astype(np.float64), g, merge=False)
    gphi, gw = a.download()
    assert (gw == ow).all() and np.array_equal(gphi, rt(ophi.astype(np.float32)))
    for h in (a, field, live, can):
        h.close()
