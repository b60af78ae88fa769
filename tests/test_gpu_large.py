"""Grids past the 32-bit fast paths: 712^3 and 713^3 voxels (361 M, 3 N >= 2^30), where the
deform call's dense sweep takes the 64-bit kernels (the flat adjoint step's 64-bit form,
aligned on 712^3 and unaligned on 713^3, and k_adj_any for the final energy; the dense scan of
the reference method; the 32-bit flat step and the nonzero masks are limited to 3 N < 2^30),
its active-set sweep lists 32-bit voxel indices with 64-bit plane offsets, and every other pass
indexes with 64-bit offsets; and 1296^3 (2.18 G voxels), whose voxel indices pass 2^31.  Each
deform runs both sweeps on the big grid.  The warp follows an active-set call (it then warps
the call's list) on every grid, and a call with another field on the grids that run both
methods (it then lists the voxels the field moves).

The scene is the sphere pair, and the big grid holds the 64^3 grid of the same scene as a
sub-box (an integer voxel offset; 2^-7 m voxels and corners on the 2^-8 m lattice, so every voxel
centre voxel (i + 1/2) + min (SDF.cpp:149-151) is exact in FP64 on both grids).  Every voxel
the frame reaches, and so every voxel that can move, lies inside the sub-box with a margin of
ten voxels; the depth frame is invalid (0) off the sphere, so no voxel outside it is observed.
Integration (SDF.cpp:137-197), both optimisers (KillingFusion.cpp:240-351), the warp
(SDF.cpp:300-328), the fusion (SDF.cpp:199-241) and marching cubes (SDF.cpp:396-414) must then
give the small grid's results bit for bit inside the sub-box and leave every other voxel as
the reset left it (phi = Max + eps, w = 0, u = 0).  The small grid's own results are pinned
against the oracle by the other GPU tests."""
import dataclasses

import numpy as np
import pytest

from killingfusion_amd import SDF, DisplacementField, KillingParams, compute_displacement_field, scenes

from helpers import sphere_frames

pytestmark = pytest.mark.gpu

V = 2.0 ** -7                      # voxel size (m)
TRUNC = 3 * V
SMALL = (64, 64, 64)
MN_S = (-0.25, -0.25, 0.546875)


class Big:
    """A big grid holding the small one at voxel offset (e, e, 0), e = (edge - 64) / 2."""

    def __init__(self, edge):
        self.dims = (edge, edge, edge)
        self.off = ((edge - SMALL[0]) // 2, (edge - SMALL[1]) // 2, 0)
        self.mn = tuple(m - o * V for m, o in zip(MN_S, self.off))

    def sub(self, a):
        """The small grid's voxels of a big-grid array (x fastest, SDF.cpp:27), in its order."""
        lead, (nx, ny, nz), (ox, oy, oz) = a.shape[:-1], self.dims, self.off
        g = a.reshape(lead + (nz, ny, nx))
        s = g[..., oz:oz + SMALL[2], oy:oy + SMALL[1], ox:ox + SMALL[0]]
        return s.reshape(lead + (-1,))


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def check_volume(B, big, small, tag):
    pb, wb = big.download()
    ps, ws = small.download()
    init = pb[0]  # big voxel 0 projects far outside the image
    assert wb[0] == 0, tag
    assert np.array_equal(B.sub(wb), ws), f"{tag}: weights differ inside the sub-box"
    assert np.array_equal(bits(B.sub(pb)), bits(ps)), f"{tag}: phi differs inside the sub-box"
    # nothing outside the sub-box changed: as many touched voxels in each grid
    assert np.count_nonzero(wb) == np.count_nonzero(ws), f"{tag}: weights outside the sub-box"
    assert np.count_nonzero(bits(pb) != bits(init)) == np.count_nonzero(bits(ps) != bits(init)), \
        f"{tag}: phi outside the sub-box"
    return int(np.count_nonzero(ws))


def check_field(B, fb, fs, tag):
    ub, us = fb.download(), fs.download()
    assert np.array_equal(bits(B.sub(ub)), bits(us)), f"{tag}: u differs inside the sub-box"
    assert np.count_nonzero(ub) == np.count_nonzero(us), f"{tag}: u nonzero outside the sub-box"
    return int(np.count_nonzero(us))


def map_to_big(B, tris):
    """Small-grid mesh coordinates (x / (dims / 2) - 1, y negated, SDF.cpp:339-347,
    SimpleMesh.h:66-80) in the big grid's normalisation."""
    p = tris.astype(np.float64).copy()
    p[..., 1] = -p[..., 1]
    for a in range(3):
        p[..., a] = ((p[..., a] + 1.0) * (SMALL[a] / 2) + B.off[a]) / (B.dims[a] / 2) - 1.0
    p[..., 1] = -p[..., 1]
    return p


@pytest.mark.parametrize("edge, dtype, methods", [
    (712, "bf16", ("adjoint", "reference")),
    (712, "f32", ("adjoint", "reference")),
    (713, "f32", ("adjoint",)),  # N % 4 = 1: the unaligned 64-bit step
    (1296, "bf16", ("adjoint",)),  # 2.18 G voxels: voxel indices past 2^31
])
def test_big_grid_equals_embedded_small_grid(ctx, edge, dtype, methods):
    B = Big(edge)
    n = edge ** 3
    assert 3 * n >= 2 ** 30
    print(f"case: {edge}^3 grid ({n / 1e6:.0f} M voxels) holding the 64^3 sphere grid, {dtype}: integrate, "
          f"deform ({', '.join(methods)}), warp, fuse, marching cubes", flush=True)
    d1, d2 = sphere_frames()
    kw = dict(voxel_size=V, clip_distance=TRUNC, max_surface=TRUNC, dtype=dtype)
    vols = {}
    for name, dims, mn in (("s", SMALL, MN_S), ("b", B.dims, B.mn)):
        can, live = SDF(ctx, dims, mn, **kw), SDF(ctx, dims, mn, **kw)
        can.integrateDepthFrame(d1, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        live.integrateDepthFrame(d2, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        vols[name] = (can, live)
    (can_s, live_s), (can_b, live_b) = vols["s"], vols["b"]
    touched = [check_volume(B, can_b, can_s, "integrate frame 1"), check_volume(B, live_b, live_s, "integrate frame 2")]
    print(f"integration: {touched} voxels observed, equal", flush=True)
    assert min(touched) > 5000

    params = {
        "adjoint": KillingParams.paper(V, 3.0, method="adjoint", max_iters=8),
        "reference": KillingParams(voxel_size=V, max_surface=TRUNC, clip_distance=TRUNC, max_iters=4),
    }
    warp_by = None
    for method in methods:
        kp = params[method]
        kd = dataclasses.replace(kp, sweep="dense")
        f_s = DisplacementField(ctx, SMALL)
        f_s.initializeAllVoxels((0.0, 0.0, 0.0))
        st_s = compute_displacement_field(live_s, can_s, f_s, kp)
        assert st_s["sweep"] == "active"
        # the dense sweep, then the automatic one (the active set: every grid of fewer than
        # 2^32 voxels can be listed), so that the warp below follows an active-set call
        for k_big, sweep in ((kd, "dense"), (kp, "active")):
            f_b = DisplacementField(ctx, B.dims)
            f_b.initializeAllVoxels((0.0, 0.0, 0.0))
            st_b = compute_displacement_field(live_b, can_b, f_b, k_big)
            tag = f"{method}, {sweep} sweep"
            assert st_b["sweep"] == sweep, tag
            for k in ("iters", "max_update", "active", "diverged_index"):
                assert st_b[k] == st_s[k], f"{tag}: {k} {st_b[k]} != {st_s[k]}"
            if method == "adjoint":  # the same per-voxel terms (all >= 0), summed in another order
                assert abs(st_b["energy"] - st_s["energy"]) <= 1e-11 * abs(st_s["energy"]), tag
            moved = check_field(B, f_b, f_s, tag)
            print(f"{tag}: {st_s['iters']} iterations, {st_s['active']} active, {st_b.get('listed')} listed, "
                  f"{moved} nonzero components, equal", flush=True)
            assert moved > 1000
            if method == "adjoint" and sweep == "active":
                warp_by = (f_s, f_b)
            else:
                f_b.close()
        if method != "adjoint":
            f_s.close()

    f_s, f_b = warp_by
    live_s.update(f_s)
    live_b.update(f_b)
    f_s.close()
    f_b.close()
    check_volume(B, live_b, live_s, "warp")
    can_s.fuse(live_s)
    can_b.fuse(live_b)
    check_volume(B, can_b, can_s, "fuse")
    print("warp and fuse: equal", flush=True)

    ms, mb = can_s.getMesh().tris, can_b.getMesh().tris
    assert ms.shape[0] > 1000 and mb.shape == ms.shape, (mb.shape, ms.shape)
    err = np.abs(mb.astype(np.float64) - map_to_big(B, ms)).max()
    print(f"marching cubes: {ms.shape[0]} triangles in both, max position difference {err:.2e}", flush=True)
    assert err <= 1.5e-7  # float32 rounding of the two normalisations
    for v in (can_s, live_s, can_b, live_b):
        v.close()
