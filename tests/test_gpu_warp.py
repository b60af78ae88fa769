"""SDF::update (kf_warp, SDF.cpp:300-328) through the C ABI against the FP64 oracle, on the
path that warps only the voxels it can change.

Where u_j = 0 the warp reads voxel j's own cell with zero weights and returns its phi and
weight unchanged (for finite phi; -0 and non-finite phi are handled by warping those voxels,
or every voxel).  A field that is zero outside a small region therefore warps in place over
the listed voxels; the result must equal the oracle's full warp bit for bit, as the dense
path's does (test_gpu_parity.py, a field that is nonzero everywhere).
"""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import DisplacementField, SDF, scenes

from helpers import C1, bf16_roundtrip, f16_roundtrip, make_pair, ogrid, sphere_frames

pytestmark = pytest.mark.gpu


def stored(phi_m, dtype, scale):
    """What a volume of this dtype returns for FP32 metres phi_m."""
    phi_m = np.asarray(phi_m, dtype=np.float32)
    if dtype == "bf16":
        return bf16_roundtrip(phi_m, scale)
    if dtype == "f16":
        return f16_roundtrip(phi_m, scale)
    return phi_m


def sparse_field(dims, seed):
    u = scenes.seeded_field(dims, amplitude=1.5, seed=seed).reshape(3, dims[2], dims[1], dims[0])
    part = np.zeros_like(u)
    part[:, 20:34, 10:30, 18:40] = u[:, 20:34, 10:30, 18:40]
    part[1, 5, 5, 5] = -0.0  # a negative zero moves nothing
    return part.reshape(3, -1)


def oracle_warp(phi, w, u):
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    ophi, ow = phi.astype(np.float64), w.astype(np.int64)
    O.warp(ophi, ow, u.astype(np.float64), g)
    return ophi, ow


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_sparse_warp_matches_oracle(ctx, dtype):
    print(f"case: 64^3 sphere frame, {dtype}, field nonzero in a 22x20x14 box, 3 voxels with phi = -0", flush=True)
    d1, d2 = sphere_frames()
    live, _ = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype=dtype)
    phi, w = live.download()
    phi = phi.copy()
    phi[[1000, 70000, 150000]] = -0.0  # u = 0 there: the warp still recomputes them
    live.upload(phi, w)
    phi, w = live.download()
    u = sparse_field(C1["dims"], seed=11)
    field = DisplacementField(ctx, C1["dims"])
    field.upload(u)
    live.update(field)
    gphi, gw = live.download()
    ophi, ow = oracle_warp(phi, w, u)
    want = stored(ophi.astype(np.float32), dtype, C1["trunc"] + 1e-5)
    moved = (np.abs(u) > 0).any(axis=0)
    print(f"moved voxels {int(moved.sum())} of {u.shape[1]}", flush=True)
    assert (gw == ow).all()
    assert np.array_equal(gphi.view(np.uint32), want.view(np.uint32))
    # voxels with u = 0 (and phi != -0) keep their phi and weight bit for bit
    keep = ~moved & (phi.view(np.uint32) != 0x80000000)
    assert np.array_equal(gphi[keep].view(np.uint32), phi[keep].view(np.uint32))
    assert (gw[keep] == w[keep]).all()


def test_sparse_warp_zero_field_is_identity(ctx):
    print("case: 64^3 sphere frame, bf16, zero field", flush=True)
    d1, d2 = sphere_frames()
    live, _ = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype="bf16")
    phi, w = live.download()
    field = DisplacementField(ctx, C1["dims"])
    live.update(field)
    gphi, gw = live.download()
    assert np.array_equal(gphi.view(np.uint32), phi.view(np.uint32)) and (gw == w).all()


def test_warp_with_a_non_finite_phi_warps_every_voxel(ctx):
    """A non-finite phi times a zero trilinear weight is NaN, so the voxels whose cell holds it
    change even where u = 0: the warp then runs over every voxel, as the oracle does."""
    print("case: 64^3 sphere frame, f32, +inf at one voxel, sparse field", flush=True)
    d1, d2 = sphere_frames()
    live, _ = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype="f32")
    phi, w = live.download()
    phi = phi.copy()
    phi[64 * 64 * 30 + 64 * 30 + 30] = np.inf
    live.upload(phi, w)
    u = sparse_field(C1["dims"], seed=12)
    field = DisplacementField(ctx, C1["dims"])
    field.upload(u)
    live.update(field)
    gphi, gw = live.download()
    ophi, ow = oracle_warp(phi, w, u)
    want = ophi.astype(np.float32)
    nan_g, nan_o = np.isnan(gphi), np.isnan(want)
    print(f"NaN voxels gpu {int(nan_g.sum())} oracle {int(nan_o.sum())}", flush=True)
    assert nan_o.sum() >= 1 and np.array_equal(nan_g, nan_o)
    assert np.array_equal(gphi[~nan_g].view(np.uint32), want[~nan_o].view(np.uint32)) and (gw == ow).all()
