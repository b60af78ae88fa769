"""The per-frame driver (kf_pipeline_process_frame, KillingFusion::processNextFrame,
KillingFusion.cpp:109-175) against the same sequence of oracle stages: integrate into a
fresh live SDF, deform onto the canonical SDF from a zero field, warp, fuse."""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import KillingFusion, KillingParams, scenes

from helpers import C1, ogrid, oparams, sphere_frames

pytestmark = pytest.mark.gpu


def oracle_sequence(frames, g, kp, adjoint, merge=True):
    trunc = C1["trunc"]
    can_phi, can_w = O.sdf_init(g)
    iters = []
    for i, d in enumerate(frames):
        if i == 0:
            O.integrate(can_phi, can_w, g, trunc, d, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
            iters.append(0)
            continue
        lphi, lw = O.sdf_init(g)
        O.integrate(lphi, lw, g, trunc, d, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        fn = O.adj_deform if adjoint else O.deform_jacobi
        u, so = fn(lphi, can_phi, np.zeros((3, g.n)), g, oparams(kp))
        iters.append(so["iters"])
        O.warp(lphi, lw, u, g)
        O.fuse(can_phi, can_w, lphi, lw, merge=merge)
    return can_phi, can_w, iters


@pytest.mark.parametrize("method", ["reference", "adjoint", "reference-nearest"])
def test_pipeline_matches_oracle_stages(ctx, method):
    print(f"case: pipeline, 3 frames of the 64^3 sphere pair, {method} method", flush=True)
    d1, d2 = sphere_frames()
    frames = [d1, d2, d1]
    merge = not method.endswith("nearest")  # FUSE_BY_MERGE (config.cpp:7)
    method = method.split("-")[0]
    if method == "reference":  # config.cpp parameters: FP32 deform parity 1e-4 voxel
        kp = KillingParams(voxel_size=C1["voxel"], max_surface=C1["trunc"], clip_distance=C1["trunc"], max_iters=16)
    else:
        kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=32)
    pl = KillingFusion(ctx, C1["dims"], C1["min3d"], C1["voxel"], kp, scenes.K, fuse_by_merge=merge)
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    stats = [pl.processNextFrame(d) for d in frames]
    gphi, gw = pl.canonical_download()
    ophi, ow, oiters = oracle_sequence(frames, g, kp, method == "adjoint", merge)
    assert [s["deform"]["iters"] for s in stats[1:]] == oiters[1:]
    touched = ow > 0
    wdiff = gw != ow
    same = ~wdiff & touched
    e = np.abs(gphi.astype(np.float64) - ophi)
    print(f"touched {touched.sum()} weight mismatches {wdiff.sum()} "
          f"phi err (equal weights): max {e[same].max():.3e} p99.9 {np.quantile(e[same], 0.999):.3e}", flush=True)
    assert touched.sum() > 10000
    assert wdiff.sum() <= 1e-3 * touched.sum()
    tol = 1e-5 if method == "reference" else 5e-5  # metres
    assert e[same].max() <= tol
    assert all(s["ms_total"] > 0 for s in stats)


def test_pipeline_frame_meshes_match_oracle(ctx):
    """The reference's three per-frame meshes (KillingFusion.cpp:157, :160, :170): the live SDF
    before the warp, after it, and the canonical SDF after the fuse, each bit-exact against the
    oracle's marching cubes on that stage's grid; on the first frame all three are the canonical
    mesh (:117-118)."""
    print("case: pipeline meshes, 3 frames of the 64^3 sphere pair, adjoint method", flush=True)
    d1, d2 = sphere_frames()
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=20)
    pl = KillingFusion(ctx, C1["dims"], C1["min3d"], C1["voxel"], kp, scenes.K)
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    trunc = C1["trunc"]

    def ref_mesh(phi):
        return O.marching_cubes(np.asarray(phi, np.float64), g, trunc, trunc).astype(np.float32)

    with pytest.raises(Exception):
        pl.frameMesh("live")  # nothing extracted yet
    pl.setMeshes()
    for k, d in enumerate([d1, d2, d1]):
        pl.processNextFrame(d)
        live_m, warped_m, can_m = pl.frameMeshes()
        cphi, _ = pl.canonical_download()
        assert np.array_equal(can_m.tris, ref_mesh(cphi)) and can_m.tris.shape[0] > 1000, k
        if k == 0:
            assert np.array_equal(live_m.tris, can_m.tris) and np.array_equal(warped_m.tris, can_m.tris)
            continue
        lphi, lw = O.sdf_init(g)
        O.integrate(lphi, lw, g, trunc, d, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        assert np.array_equal(live_m.tris, ref_mesh(lphi.astype(np.float32))) and live_m.tris.shape[0] > 1000, k
        wphi, _ = pl.live_download()
        assert np.array_equal(warped_m.tris, ref_mesh(wphi)), k
        assert not np.array_equal(warped_m.tris, live_m.tris)  # the warp moved the surface
    pl.setMeshes(("canonical",))
    pl.processNextFrame(d2)
    assert pl.frameMesh("canonical").tris.shape[0] > 1000
    with pytest.raises(Exception):
        pl.frameMesh("live")
    pl.close()


def test_pipeline_field_dumps(ctx, tmp_path):
    """The reference writes each frame's displacement field after its deform
    (KillingFusion.cpp:146-151), named by the dataset frame number; frame 0 has no deform."""
    print("case: pipeline per-frame field dumps, 3 frames of the 64^3 sphere pair", flush=True)
    from killingfusion_amd import formats
    d1, d2 = sphere_frames()
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=20)
    pl = KillingFusion(ctx, C1["dims"], C1["min3d"], C1["voxel"], kp, scenes.K)
    pl.setFieldDumpPrefix(tmp_path / "frame_", start_frame=5, stride=3)
    fields = []
    for d in (d1, d2, d1):
        pl.processNextFrame(d)
        fields.append(pl.field_download())
    assert not (tmp_path / "frame_005.bin").exists()
    for k, name in ((1, "frame_008.bin"), (2, "frame_011.bin")):
        u = formats.read_displacement_bin(tmp_path / name)
        assert u["dims"] == tuple(C1["dims"]) and u["voxel"] == C1["voxel"]
        assert np.array_equal(u["u"], fields[k].astype(np.float64)), name
    assert np.abs(fields[1]).max() > 0
    pl.close()
