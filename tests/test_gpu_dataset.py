"""The dataset front end into the per-frame driver on the GPU: a synthetic copy of the dataset
layout (three hinge frames as 16-bit millimetre PNGs with object masks), the reference's grid
over the camera frustum (KillingFusion.cpp:10-31, 478-519), and every frame of the driver
(KillingFusion::processNextFrame, KillingFusion.cpp:109-175) checked stage by stage against the
oracle on the same masked depth: frame 0's integration bit for bit; on frames 1 and 2 the
integration of the live frame, the warp and the fusion bit for bit, and the deformation to
one-step parity from the GPU's own field (the two FP32/FP64 trajectories part at trilinear
cell-face ties, so the whole-call result is held to the 1e-3-voxel bar instead)."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from killingfusion_amd import DisplacementField, KillingParams, SDF, compute_displacement_field, dataset, scenes

from helpers import oparams
from test_dataset import make_dataset

pytestmark = pytest.mark.gpu


def test_dataset_frames_through_the_frame_driver(ctx, tmp_path):
    print("case: dataset reader -> frustum grid -> 3 frames through kf_pipeline", flush=True)
    make_dataset(tmp_path, [scenes.hinge_frame_exact(t, frame=i) for i, t in enumerate((0.0, 0.3, 0.6))])
    r = dataset.DatasetReader(tmp_path, "Snoopy")
    kp = KillingParams(method="adjoint", max_iters=64)  # config.cpp values: 12 mm voxels, Max = clip = 4 voxels
    pl, dims, mn = dataset.pipeline_for(r, ctx, kp)
    print("grid", dims, mn, flush=True)
    assert dims[2] == int((1.2 - 0.0495164) / 0.012) + 1
    depth0 = r.depth_frame(0)
    st0 = pl.processNextFrame(depth0)
    gphi, gw = pl.canonical_download()
    g = O.OGrid(dims=dims, min_corner=mn, voxel=kp.voxel_size, oob=kp.max_surface + kp.epsilon)
    ophi, ow = O.sdf_init(g)
    O.integrate(ophi, ow, g, kp.clip_distance, depth0, r.K, np.eye(4), r.min_depth, r.max_depth)
    assert (gw == ow).all() and np.array_equal(gphi, ophi.astype(np.float32)) and (gw > 0).sum() > 5000
    clip, vox = kp.clip_distance, kp.voxel_size
    for i in (1, 2):
        print(f"case: dataset frame {i} stages against the oracle", flush=True)
        cphi0, cw0 = pl.canonical_download()
        depth = r.depth_frame(i)
        st = pl.processNextFrame(depth)
        assert st["frame"] == i and st["deform"]["iters"] == 64 and st["deform"]["diverged_index"] == -1
        assert st["deform"]["active"] > 0
        # integration of the live frame (FP32 storage: the stored grid is the FP64 result in FP32)
        lphi_o, lw_o = O.sdf_init(g)
        O.integrate(lphi_o, lw_o, g, clip, depth, r.K, np.eye(4), r.min_depth, r.max_depth)
        lstored = lphi_o.astype(np.float32)
        # deformation of the live grid onto the previous canonical one
        uo, so = O.adj_deform(lstored.astype(np.float64), cphi0.astype(np.float64), np.zeros((3, g.n)), g,
                              oparams(kp))
        ug = pl.field_download()
        moved = np.any(uo != 0, axis=0) | np.any(ug != 0, axis=0)
        e = np.abs(ug.astype(np.float64) - uo).max(axis=0)[moved]
        sl, sc = (SDF(ctx, dims, mn, voxel_size=vox, clip_distance=clip, max_surface=kp.max_surface)
                  for _ in range(2))
        sl.upload(lstored, lw_o.astype(np.int32))
        sc.upload(cphi0, cw0)
        f = DisplacementField(ctx, dims)
        f.upload(ug)
        one = dataclasses.replace(kp, max_iters=1)
        compute_displacement_field(sl, sc, f, one)
        step_o, _ = O.adj_deform(lstored.astype(np.float64), cphi0.astype(np.float64), ug.astype(np.float64), g,
                                 oparams(one))
        step_e = float(np.abs(f.download().astype(np.float64) - step_o).max())
        for h in (f, sl, sc):
            h.close()
        print(f"frame {i}: moved {moved.sum()} trajectory max {e.max():.3e} n>1e-6 {(e > 1e-6).sum()} | "
              f"one step from the GPU's field: max {step_e:.3e} voxel", flush=True)
        assert so["iters"] == 64 and moved.sum() > 100
        assert step_e <= 1e-6
        assert e.max() <= 1e-3 and (e > 1e-6).sum() <= 0.01 * e.size
        # warp of the integrated live grid by the GPU's field, then fusion into the previous canonical
        wphi, ww = lstored.astype(np.float64), lw_o.copy()
        O.warp(wphi, ww, ug.astype(np.float64), g)
        gl_phi, gl_w = pl.live_download()
        assert (gl_w == ww).all() and np.array_equal(gl_phi, wphi.astype(np.float32)), f"frame {i}: warp"
        fphi, fw = cphi0.astype(np.float64), cw0.astype(np.int64)
        O.fuse(fphi, fw, gl_phi.astype(np.float64), gl_w.astype(np.int64))
        cphi, cw = pl.canonical_download()
        assert (cw == fw).all() and np.array_equal(cphi, fphi.astype(np.float32)), f"frame {i}: fuse"
    assert st0["frame"] == 0
    assert pl.canonical_mesh().tris.shape[0] > 1000
