"""The dataset front end into the per-frame driver on the GPU: a synthetic copy of the dataset
layout (three hinge frames as 16-bit millimetre PNGs with object masks), the reference's grid
over the camera frustum (KillingFusion.cpp:10-31, 478-519), and the driver's first frame
checked bit for bit against the oracle's integration of the same masked depth."""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import KillingParams, dataset, scenes

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
    for i in (1, 2):
        st = pl.processNextFrame(r.depth_frame(i))
        assert st["frame"] == i and st["deform"]["iters"] == 64 and st["deform"]["diverged_index"] == -1
        assert st["deform"]["active"] > 0
    assert st0["frame"] == 0
    assert pl.canonical_mesh().tris.shape[0] > 1000
