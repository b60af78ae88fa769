"""Writes tests/golden/sphere16.npz: fixed-scene vectors from the FP64 oracle for regression
checks of both the oracle (tests/test_golden.py) and the HIP library (tests/test_gpu_golden.py).

Scene: the sphere pair of the 64^3 workload (sphere r = 0.15 m at 0.8 m, frame 2 with a 1 cm
radial bulge), integrated into a 16^3 grid of 32 mm voxels, min corner (-0.256, -0.256, 0.544),
truncation 3 voxels.  Contents: integrated phi and weights of both frames; 8 Jacobi iterations
from a zero field of the reference formula (config.cpp:45-54 parameters: alpha 0.025,
omega_k 0.04, gamma 0, omega_s 0.1) and of the adjoint formulation (paper parameters: alpha 0.1,
omega_k 0.5, gamma 0.1, omega_s 0.2); the marching-cubes soup of the live SDF.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

DIMS, MIN3D, VOXEL, TRUNC = (16, 16, 16), (-0.256, -0.256, 0.544), 0.032, 0.096
ITERS = 8


def compute():
    import oracle as O
    from killingfusion_amd import scenes
    d1, d2 = scenes.sphere_pair()
    g = O.OGrid(dims=DIMS, min_corner=MIN3D, voxel=VOXEL, oob=TRUNC + 1e-5)
    out = {}
    for name, d in (("can", d1), ("live", d2)):
        phi, w = O.sdf_init(g)
        O.integrate(phi, w, g, TRUNC, d, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
        out[f"{name}_phi"], out[f"{name}_w"] = phi, w
    p_ref = O.OParams(voxel_size=VOXEL, max_surface=TRUNC, clip=TRUNC, max_iters=ITERS)
    p_adj = O.OParams(voxel_size=VOXEL, max_surface=TRUNC, clip=TRUNC, alpha=0.1, omega_k=0.5, gamma=0.1,
                      omega_s=0.2, max_iters=ITERS)
    lphi32 = out["live_phi"].astype(np.float32).astype(np.float64)
    cphi32 = out["can_phi"].astype(np.float32).astype(np.float64)
    out["u_reference"], _ = O.deform_jacobi(lphi32, cphi32, np.zeros((3, g.n)), g, p_ref)
    out["u_adjoint"], _ = O.adj_deform(lphi32, cphi32, np.zeros((3, g.n)), g, p_adj)
    out["mc_live"] = O.marching_cubes(lphi32, g, TRUNC, TRUNC)
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "sphere16.npz"), **compute())
    print("wrote", os.path.join(HERE, "sphere16.npz"))
