"""Welded marching cubes (kf_marching_cubes_indexed).  Its triangles are the soup's triangles
(kf_marching_cubes, bit-exact against the oracle) in the same order, with each vertex shared
by every triangle that cuts the same grid edge; on a closed surface the result is a closed
2-manifold, and the normals follow the SDF gradient."""
import numpy as np
import pytest

from killingfusion_amd import SDF

from helpers import C1

pytestmark = pytest.mark.gpu


def test_welded_mesh_is_the_soup_indexed(ctx):
    print("case: welded mesh vs soup, 64^3 sphere frame", flush=True)
    from helpers import make_pair, sphere_frames
    d1, d2 = sphere_frames()
    live, _ = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1)
    soup = live.getMesh().tris.astype(np.float64)
    m = live.getMeshIndexed()
    print("soup tris", soup.shape[0], "welded verts", m.vertices.shape[0], flush=True)
    assert m.triangles.shape[0] == soup.shape[0] > 1000
    assert m.triangles.max() < m.vertices.shape[0]
    # every vertex is used, and is the soup's vertex up to the orientation of the edge it lies on
    assert np.unique(m.triangles).size == m.vertices.shape[0]
    assert np.abs(m.vertices[m.triangles].astype(np.float64) - soup).max() <= 1e-6
    # welding merges each cut edge's copies: far fewer vertices than the soup's 3T
    assert m.vertices.shape[0] < 0.6 * 3 * soup.shape[0]
    assert np.allclose(np.linalg.norm(m.normals, axis=1), 1.0, atol=1e-5)


def test_welded_full_sphere_is_closed_with_outward_normals(ctx):
    print("case: welded mesh of an analytic full sphere, 48^3", flush=True)
    n, vox, trunc = 48, 0.01, 0.05
    c = np.array([0.231, 0.244, 0.237])
    r = 0.15
    idx = np.arange(n) * vox + 0.5 * vox
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    phi = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r
    phi = np.clip(phi, -trunc, trunc).astype(np.float32).ravel()
    sdf = SDF(ctx, (n, n, n), (0, 0, 0), vox, trunc, trunc)
    sdf.upload(phi, np.ones(n ** 3, dtype=np.int32))
    m = sdf.getMeshIndexed()
    V, F = m.vertices.shape[0], m.triangles.shape[0]
    t = m.triangles.astype(np.int64)
    edges = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    uniq, counts = np.unique(edges, axis=0, return_counts=True)
    print(f"V {V} F {F} E {uniq.shape[0]} chi {V - uniq.shape[0] + F}", flush=True)
    assert (counts == 2).all()          # every edge between exactly two triangles
    assert V - uniq.shape[0] + F == 2   # a sphere
    # normals point away from the centre, nearly radially (normalised coordinates are isotropic here)
    centre = m.vertices.mean(axis=0)
    radial = m.vertices - centre
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    cosang = np.sum(radial * m.normals, axis=1)
    assert cosang.min() > 0.98
