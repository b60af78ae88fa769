"""The reference optimiser's trust region (UseTrustStrategy, KillingFusion.cpp:288-314) on the GPU
against the FP64 oracle, through the C ABI.

With the data energy alone, each voxel's step -alpha g is shrunk (alpha / 1.5 each time) until
the data residual at the moved sample is no larger than at the current one; a voxel whose
alpha falls below 1e-7 takes no step, and the reference's `break` leaves the rest of its x-row
without a step in that iteration.  The GPU takes these decisions in FP64 like the reference
(its gradient is the FP32 one of the reference-formula path), then a pass over the band list
undoes the steps past each row's first failure.  Random fields make both the shrinking and the
row breaks frequent.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from killingfusion_amd import DisplacementField, KillingParams, SDF, compute_displacement_field
from killingfusion_amd._lib import KFError, KF_E_INVALID_ARG

from helpers import ogrid, oparams

pytestmark = pytest.mark.gpu

N = 24
VOX = 0.01
TRUNC = 0.03


@pytest.fixture(scope="module")
def noisy(ctx):
    rng = np.random.default_rng(5)
    n = N ** 3
    live = rng.uniform(-0.03, 0.03, n).astype(np.float32)
    can = rng.uniform(-0.03, 0.03, n).astype(np.float32)
    w = np.ones(n, dtype=np.int32)
    vols = []
    for phi in (live, can):
        v = SDF(ctx, (N, N, N), (0.0, 0.0, 0.0), VOX, TRUNC, TRUNC)
        v.upload(phi, w)
        vols.append(v)
    return vols[0], vols[1], live, can


def params(**kw):
    base = dict(voxel_size=VOX, max_surface=TRUNC, clip_distance=TRUNC, alpha=0.1, use_levelset=False,
                use_killing=False, use_trust=True, max_iters=3)
    base.update(kw)
    return KillingParams(**base)


def run(ctx, live, can, kp, sweep):
    f = DisplacementField(ctx, (N, N, N))
    try:
        st = compute_displacement_field(live, can, f, dataclasses.replace(kp, sweep=sweep))
        return f.download(), st
    finally:
        f.close()


@pytest.mark.parametrize("iters", [1, 3])
def test_trust_region_matches_oracle(ctx, noisy, iters):
    live, can, lphi, cphi = noisy
    kp = params(max_iters=iters)
    print(f"case: trust region, random {N}^3 fields, data energy alone, {iters} iterations", flush=True)
    g = ogrid((N, N, N), (0.0, 0.0, 0.0), VOX, TRUNC)
    uo, so = O.deform_jacobi(lphi.astype(np.float64), cphi.astype(np.float64), np.zeros((3, N ** 3)), g,
                             oparams(kp))
    plain, _ = O.deform_jacobi(lphi.astype(np.float64), cphi.astype(np.float64), np.zeros((3, N ** 3)), g,
                               oparams(dataclasses.replace(kp, use_trust=False)))
    ug, st = run(ctx, live, can, kp, "dense")
    e = np.abs(ug.astype(np.float64) - uo).max(axis=0)
    moved_o = np.any(uo != 0, axis=0)
    moved_g = np.any(ug != 0, axis=0)
    print(f"iters {st['iters']}/{so['iters']} active {st['active']}/{so['active']} "
          f"moved {moved_g.sum()}/{moved_o.sum()} (plain step moves {np.any(plain != 0, axis=0).sum()}) "
          f"max {e.max():.3e} n>1e-5 {(e > 1e-5).sum()} moved-set mismatches {(moved_o != moved_g).sum()}",
          flush=True)
    assert st["iters"] == so["iters"] == iters
    assert np.any(plain != uo), "the trust region changes the result on these fields"
    assert (moved_o != moved_g).sum() <= 2
    assert (e > 1e-5).sum() <= 2 and np.quantile(e, 0.999) <= 1e-6
    if iters == 1:
        assert st["active"] == so["active"]


def test_trust_region_active_equals_dense(ctx, noisy):
    live, can, _, _ = noisy
    kp = params()
    print("case: trust region, dense vs active sweep", flush=True)
    a, b = run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active")
    assert a[1]["sweep"] == "dense" and b[1]["sweep"] == "active"
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    for k in ("iters", "max_update", "active", "diverged_index"):
        assert a[1][k] == b[1][k], k


def test_trust_flag_needs_the_data_energy_alone(ctx, noisy):
    """The reference honours UseTrustStrategy only with the data energy alone (:288); with
    another energy on the flag changes nothing.  The adjoint method has no trust region."""
    live, can, _, _ = noisy
    print("case: trust flag with the level-set energy on; with the adjoint method", flush=True)
    on = params(use_levelset=True)
    a, b = run(ctx, live, can, on, "dense"), run(ctx, live, can, dataclasses.replace(on, use_trust=False), "dense")
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    f = DisplacementField(ctx, (N, N, N))
    with pytest.raises(KFError) as err:
        compute_displacement_field(live, can, f, params(method="adjoint"))
    assert err.value.status == KF_E_INVALID_ARG
    f.close()
