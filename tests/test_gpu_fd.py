"""Single frame pair at 128^3 (8 mm voxels, the sphere pair), u seeded U(-0.5, 0.5) voxel: the
fused adjoint kernel's gradient (kf_energy_backward) against FP64 central differences of the
oracle energy (eps 1e-3 voxel) at 4096 random active voxels, and the plain reference-formula
gradient against the oracle's."""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import DisplacementField, KillingParams, energy_backward, energy_gradient_ref, scenes

from helpers import make_pair, ogrid, oparams, sphere_frames

pytestmark = pytest.mark.gpu

DIMS = (128, 128, 128)
MIN3D = (-0.512, -0.512, 0.288)
VOXEL, TRUNC = 0.008, 0.024


@pytest.fixture(scope="module")
def pair128(ctx):
    d1, d2 = sphere_frames()
    live, can = make_pair(ctx, DIMS, MIN3D, VOXEL, TRUNC, d2, d1)
    lphi, _ = live.download()
    cphi, _ = can.download()
    u = scenes.seeded_field(DIMS, amplitude=0.5, seed=0x5EED)
    field = DisplacementField(ctx, DIMS)
    field.upload(u)
    return dict(live=live, can=can, lphi=lphi.astype(np.float64), cphi=cphi.astype(np.float64),
                u=u.astype(np.float64), field=field, g=ogrid(DIMS, MIN3D, VOXEL, TRUNC))


def test_adjoint_gradient_vs_fp64_central_differences(ctx, pair128):
    print("case: 128^3 adjoint gradient vs FP64 central differences, 4096 active voxels", flush=True)
    P = pair128
    kp = KillingParams.paper(VOXEL, 3.0, method="adjoint")
    p = oparams(kp)
    out = DisplacementField(ctx, DIMS)
    energy_backward(P["live"], P["can"], P["field"], kp, out)
    gg = out.download().astype(np.float64)
    _, mask = O.adj_gradient(P["lphi"], P["cphi"], P["u"], P["g"], p)
    act = np.nonzero(mask)[0]
    assert act.size > 5000
    rng = np.random.default_rng(2024)
    picks = rng.choice(act, 4096, replace=False)
    eps = 1e-3
    u = np.ascontiguousarray(P["u"])
    nx, ny = DIMS[0], DIMS[1]
    errs, checked, skipped = [], 0, 0
    for i in picks:
        xyz = (int(i % nx), int((i // nx) % ny), int(i // (nx * ny)))
        for a in range(3):
            q = xyz[a] + u[a, i]
            if abs((q - np.floor(q)) - 0.5) > 0.5 - 2 * eps:  # the +-eps moves would cross a cell face
                skipped += 1
                continue
            u0 = u[a, i]
            u[a, i] = u0 + eps
            ep = O.adj_energy_local(P["lphi"], P["cphi"], u, P["g"], p, *xyz)
            u[a, i] = u0 - eps
            em = O.adj_energy_local(P["lphi"], P["cphi"], u, P["g"], p, *xyz)
            u[a, i] = u0
            fd = (ep - em) / (2 * eps)
            errs.append(abs(fd - gg[a, i]) / (1.0 + abs(fd)))
            checked += 1
    errs = np.array(errs)
    bad = int((errs > 1e-4).sum())
    print(f"checked {checked} components (skipped {skipped} at cell faces): max rel err {errs.max():.3e}, "
          f"p99.9 {np.quantile(errs, 0.999):.3e}, above 1e-4: {bad}", flush=True)
    assert checked > 10000
    # a component whose +-eps move flips the voxel's band mask is the only allowed outlier
    assert bad <= 0.001 * checked
    assert np.quantile(errs, 0.999) <= 1e-4


def test_reference_gradient_128_matches_oracle(ctx, pair128):
    print("case: 128^3 reference-formula gradient vs oracle", flush=True)
    P = pair128
    kp = KillingParams.paper(VOXEL, 3.0)
    out = DisplacementField(ctx, DIMS)
    energy_gradient_ref(P["live"], P["can"], P["field"], kp, out)
    gg = out.download().astype(np.float64)
    og = O.energy_gradient_grid(P["lphi"], P["cphi"], P["u"], P["g"], oparams(kp), masked=True)
    gm, om = np.any(gg != 0, axis=0), np.any(og != 0, axis=0)
    both = gm & om
    print(f"mask mismatches {(gm != om).sum()} of {om.sum()}; max err "
          f"{np.abs(gg[:, both] - og[:, both]).max():.3e}", flush=True)
    assert (gm != om).sum() <= 0.001 * om.sum()
    assert both.sum() > 5000
    assert np.abs(gg[:, both] - og[:, both]).max() <= 1e-4 * (1 + np.abs(og[:, both]).max())
