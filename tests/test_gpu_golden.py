"""The HIP library against the committed fixed-scene vectors (tests/golden/sphere16.npz):
integration and marching cubes bit for bit, 8 deform iterations of each method within 1e-4
voxel (the reference formula with its config.cpp parameters, the adjoint with the paper's)."""
import os

import numpy as np
import pytest

from killingfusion_amd import DisplacementField, KillingParams, SDF, compute_displacement_field, scenes

from test_golden import load_maker

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_gpu_matches_golden_vectors(ctx):
    print("case: 16^3 golden sphere vectors", flush=True)
    m = load_maker()
    gold = np.load(os.path.join(HERE, "golden", "sphere16.npz"))
    d1, d2 = scenes.sphere_pair()
    kw = dict(voxel_size=m.VOXEL, clip_distance=m.TRUNC, max_surface=m.TRUNC)
    vols = {}
    for name, d in (("can", d1), ("live", d2)):
        v = SDF(ctx, m.DIMS, m.MIN3D, **kw)
        v.integrateDepthFrame(d, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        phi, w = v.download()
        assert np.array_equal(phi, gold[f"{name}_phi"].astype(np.float32)) and (w == gold[f"{name}_w"]).all(), name
        vols[name] = v
    tris = vols["live"].getMesh().tris
    assert np.array_equal(tris, gold["mc_live"].astype(np.float32))
    for method in ("reference", "adjoint"):
        if method == "reference":  # config.cpp parameters
            kp = KillingParams(voxel_size=m.VOXEL, max_surface=m.TRUNC, clip_distance=m.TRUNC, max_iters=m.ITERS)
        else:
            kp = KillingParams.paper(m.VOXEL, 3.0, method=method, max_iters=m.ITERS)
        f = DisplacementField(ctx, m.DIMS)
        st = compute_displacement_field(vols["live"], vols["can"], f, kp)
        err = np.abs(f.download().astype(np.float64) - gold[f"u_{method}"]).max()
        print(method, "max err", err, flush=True)
        assert st["iters"] == m.ITERS and err <= 1e-4
