"""Marching cubes (SDF::getMesh, SDF.cpp:330-414; Polygonise, MarchingCubes.h:326-422) on grids
whose rows are not a multiple of the cells a thread classifies, and the count cache.

A thread of the count pass classifies eight consecutive cells: on grids whose rows are a
multiple of 8 nodes, from four aligned node rows with the thresholds mapped to raw stored
values; otherwise from four node rows when the cells and their x + 1 corners share one row,
cell by cell where they wrap into the next row or run past the grid.  The emit pass classifies
again, one cell per thread.  Every path must give the oracle's soup bit for bit.  A size query
followed by the extraction counts the volume once; any change to the volume (integrate, fuse,
warp, upload, copy, reset) or another iso value must count again.
"""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import SDF, DisplacementField, scenes

from helpers import ogrid

pytestmark = pytest.mark.gpu

TRUNC = 0.03
VOXEL = 0.01


def wavy(dims, seed):
    """A smooth field with many zero crossings, in [-0.05, 0.05] m (cells beyond the band skip)."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.random.default_rng(seed).uniform(0.2, 0.6, 6)
    f = 0.05 * np.sin(r[0] * x + r[3]) * np.cos(r[1] * y + r[4]) * np.sin(r[2] * z + r[5])
    return f.astype(np.float32).ravel()


def soup_ref(vol, iso=0.0):
    phi, _ = vol.download()
    g = ogrid(vol.dims, vol.min3d, VOXEL, TRUNC)
    return O.marching_cubes(phi.astype(np.float64), g, TRUNC, TRUNC, iso).astype(np.float32)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("dims", [(61, 45, 58), (70, 45, 58), (5, 9, 7), (64, 33, 20), (4, 4, 4)])
def test_marching_cubes_uneven_rows(ctx, dims, dtype):
    print(f"case: marching cubes, wavy field {dims}, {dtype}", flush=True)
    v = SDF(ctx, dims, (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC, dtype=dtype)
    v.upload(wavy(dims, sum(dims)))
    ref = soup_ref(v)
    mesh = v.getMesh()
    assert mesh.tris.shape[0] == ref.shape[0]
    assert np.array_equal(mesh.tris, ref)
    if min(dims) > 8:
        assert ref.shape[0] > 1000
    v.close()


def test_marching_cubes_count_cache(ctx):
    """Alternating volumes, iso values and changes between the size query and the extraction."""
    print("case: marching cubes count cache, 40^3 and 33^3 volumes", flush=True)
    a = SDF(ctx, (40, 40, 40), (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC)
    b = SDF(ctx, (33, 33, 33), (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC)
    a.upload(wavy(a.dims, 1))
    b.upload(wavy(b.dims, 2))
    ra, rb = soup_ref(a), soup_ref(b)
    for _ in range(2):
        assert np.array_equal(a.getMesh().tris, ra)
        assert np.array_equal(b.getMesh().tris, rb)
    assert np.array_equal(a.getMesh(0.01).tris, soup_ref(a, 0.01))
    assert np.array_equal(a.getMesh().tris, ra)
    # each kind of change to a after its count
    c = SDF(ctx, (40, 40, 40), (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC)
    c.upload(wavy(a.dims, 3), np.ones(a.n, dtype=np.int32))
    field = DisplacementField(ctx, a.dims)
    field.upload(scenes.seeded_field(a.dims, amplitude=0.7, seed=4))
    changes = [
        ("upload", lambda: a.upload(wavy(a.dims, 5), np.ones(a.n, dtype=np.int32))),
        ("fuse", lambda: a.fuse(c)),
        ("warp", lambda: a.update(field)),
        ("copy", lambda: a.copy_from(c)),
        ("integrate", lambda: a.integrateDepthFrame(scenes.sphere_pair()[0], scenes.C2W, scenes.K, 0.1, 3.0)),
        ("reset", lambda: a.reset()),
    ]
    for name, change in changes:
        before = a.getMesh().tris
        change()
        after = a.getMesh().tris
        ref = soup_ref(a)
        print(f"{name}: {before.shape[0]} -> {after.shape[0]} triangles", flush=True)
        assert np.array_equal(after, ref), name
    # a volume created where a destroyed one lived counts afresh
    b.close()
    d = SDF(ctx, (33, 33, 33), (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC)
    d.upload(-wavy(d.dims, 2))
    assert np.array_equal(d.getMesh().tris, soup_ref(d))
    for h in (a, c, d, field):
        h.close()


def threshold_neighbours(t):
    """The float nearest t and the floats on either side of it."""
    f = np.float32(t)
    return [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]


@pytest.mark.parametrize("dims", [(23, 17, 19), (24, 17, 19)])
@pytest.mark.parametrize("iso", [0.0123456789, 0.0, 0.05, -0.0299999])
def test_marching_cubes_threshold_edges(ctx, iso, dims):
    """The passes compare corner values in FP32 against thresholds rounded outward from the
    FP64 ones (iso, MaxSurfaceVoxelDistance, -UnknownClipDistance; none of them a float here
    but 0): values drawn from the floats at and next to every threshold must classify, skip
    and interpolate exactly as the oracle's FP64 comparisons do.  iso = 0.05 lies above the
    out-of-grid value (Max + eps), so boundary corners count as below the surface.  Rows of 24
    nodes take the count pass's aligned path, rows of 23 its general one."""
    print(f"case: marching cubes at threshold-adjacent values, {dims}, iso {iso}", flush=True)
    pool = []
    for t in (iso, TRUNC, -TRUNC, TRUNC + 1e-5):
        pool += threshold_neighbours(t)
    rng = np.random.default_rng(17)
    pool = np.array(pool + list(rng.uniform(-0.06, 0.06, 24)), dtype=np.float32)
    n = dims[0] * dims[1] * dims[2]
    phi = rng.choice(pool, n).astype(np.float32)
    v = SDF(ctx, dims, (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC, dtype="f32")
    v.upload(phi)
    g = ogrid(dims, (0.0, 0.0, 0.0), VOXEL, TRUNC)
    ref = O.marching_cubes(phi.astype(np.float64), g, TRUNC, TRUNC, iso).astype(np.float32)
    mesh = v.getMesh(iso)
    assert ref.shape[0] > 500
    assert mesh.tris.shape[0] == ref.shape[0]
    assert np.array_equal(mesh.tris, ref)
    v.close()


def _key16(c):
    """16-bit sign-magnitude codes (BF16, FP16) in increasing order of their values."""
    return (~c & 0xFFFF) if c & 0x8000 else (c | 0x8000)


def _unkey16(k):
    return (k & 0x7FFF) if k & 0x8000 else (~k & 0xFFFF)


def stored_neighbours(t, dtype, scale):
    """Decoded values of the five stored codes nearest t / scale (a BF16 or FP16 code decodes
    to float32(code) * scale in float32), so that they straddle the decoded threshold t."""
    x = np.float32(np.float32(t) / scale)
    if dtype == "bf16":
        c = int(np.array([x], np.float32).view(np.uint32)[0]) >> 16
        codes = [_unkey16(_key16(c) + k) for k in range(-2, 3)]
        raw = (np.array(codes, np.uint32) << 16).view(np.float32)
    else:
        c = int(np.array([x], np.float16).view(np.uint16)[0])
        codes = [_unkey16(_key16(c) + k) for k in range(-2, 3)]
        raw = np.array(codes, np.uint16).view(np.float16).astype(np.float32)
    return list(raw * np.float32(scale))


@pytest.mark.parametrize("dims", [(23, 17, 19), (24, 17, 19)])
@pytest.mark.parametrize("iso", [0.0123456789, 0.0, 0.05])
@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_marching_cubes_threshold_edges_stored(ctx, iso, dims, dtype):
    """BF16 and FP16 volumes whose stored codes sit at and next to every threshold (iso,
    MaxSurfaceVoxelDistance, -UnknownClipDistance, the out-of-grid value): the aligned count
    path compares raw codes against thresholds mapped through the decode, the general path and
    the emit pass compare decoded values, and all must classify as the oracle does on the
    downloaded grid."""
    print(f"case: marching cubes at stored threshold-adjacent codes, {dims}, {dtype}, iso {iso}", flush=True)
    scale = np.float32(TRUNC + 1e-5)  # the decode scale: Max + eps, as a float
    pool = []
    for t in (iso, TRUNC, -TRUNC, TRUNC + 1e-5):
        pool += stored_neighbours(t, dtype, scale)
    rng = np.random.default_rng(23)
    pool = np.array(pool + list(rng.uniform(-0.06, 0.06, 24)), dtype=np.float32)
    n = dims[0] * dims[1] * dims[2]
    v = SDF(ctx, dims, (0.0, 0.0, 0.0), VOXEL, TRUNC, TRUNC, dtype=dtype)
    v.upload(rng.choice(pool, n).astype(np.float32))
    phi, _ = v.download()
    g = ogrid(dims, (0.0, 0.0, 0.0), VOXEL, TRUNC)
    ref = O.marching_cubes(phi.astype(np.float64), g, TRUNC, TRUNC, iso).astype(np.float32)
    mesh = v.getMesh(iso)
    assert ref.shape[0] > 500
    assert mesh.tris.shape[0] == ref.shape[0]
    assert np.array_equal(mesh.tris, ref)
    v.close()
