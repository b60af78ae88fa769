"""Batched deform (kf_batch_deform / kf_batch_autotune) through the C ABI.

The batched launch runs the same per-voxel arithmetic as kf_deform, so every volume's field
and stats must equal a per-volume kf_deform run bit for bit (kf_deform itself is held to the
FP64 oracle in test_gpu_parity.py).  The iteration loop is a captured hipGraph: each test
replays it several times with other GPU work in between and checks every replay.
"""
import numpy as np
import pytest

from killingfusion_amd import (SDF, DisplacementField, KillingParams, batch_autotune, batch_deform,
                               compute_displacement_field, scenes)
from killingfusion_amd._lib import KF_E_DIVERGED, KF_E_INVALID_ARG, KFError

from helpers import listed_on_cpu, make_pair

pytestmark = pytest.mark.gpu

VOX = scenes.BATCH_VOXEL
TRUNC = 3 * VOX


@pytest.fixture(scope="module")
def batch(ctx):
    """Three cubes (48, 61, 77: odd and even sizes, none a multiple of a tile) and one non-cubic
    volume."""
    print("case: batch of 4 ellipsoid volumes (48^3, 61^3, 77^3, 70x45x58), 6 mm", flush=True)
    vols = []
    for n, mn, d1, d2 in scenes.batch_scene(3, seed=0x5EED, edges=[48, 61, 77]):
        vols.append(make_pair(ctx, (n, n, n), mn, VOX, TRUNC, d2, d1))
    dims = (70, 45, 58)
    ext = [d * VOX for d in dims]
    d1, d2, _ = scenes.ellipsoid_pair(0x5EED, 3, max_axis=0.4 * min(ext))
    vols.append(make_pair(ctx, dims, (-ext[0] / 2, -ext[1] / 2, 0.8 - ext[2] / 2), VOX, TRUNC, d2, d1))
    return vols


def params(sweep="dense", **kw):
    return KillingParams.paper(VOX, 3.0, method="adjoint", sweep=sweep, **kw)


def per_volume(ctx, vols, kp, inits=None):
    out, st = [], []
    for b, (live, can) in enumerate(vols):
        f = DisplacementField(ctx, live.dims)
        if inits is not None:
            f.upload(inits[b])
        st.append(compute_displacement_field(live, can, f, kp))
        out.append(f.download())
    return out, st


def fresh_fields(ctx, vols, inits=None):
    fs = []
    for b, (live, _) in enumerate(vols):
        f = DisplacementField(ctx, live.dims)
        if inits is not None:
            f.upload(inits[b])
        fs.append(f)
    return fs


def lists(vols):
    return [v[0] for v in vols], [v[1] for v in vols]


@pytest.mark.parametrize("capture", [False, True])
@pytest.mark.parametrize("sweep", ["dense", "active"])
def test_batch_equals_per_volume_across_replays(ctx, batch, sweep, capture):
    print(f"case: batch of 4, 50 iterations, {sweep} sweep, loop capture {capture}", flush=True)
    ctx.set_loop_capture(capture)
    try:
        _batch_across_replays(ctx, batch, sweep)
    finally:
        ctx.set_loop_capture(False)


def _batch_across_replays(ctx, batch, sweep):
    kp = params(sweep, max_iters=50)
    ref, ref_st = per_volume(ctx, batch, kp)
    assert all(s["active"] > 0 for s in ref_st)
    lives, cans = lists(batch)
    fields = fresh_fields(ctx, batch)
    for rep in range(3):
        print(f"replay {rep}", flush=True)
        for f in fields:
            f.initializeAllVoxels((0.0, 0.0, 0.0))
        st = batch_deform(lives, cans, fields, kp, stats=(rep != 1))
        for b, f in enumerate(fields):
            assert np.array_equal(f.download(), ref[b]), (rep, b)
        if st is not None:
            for b in range(len(batch)):
                assert st[b]["sweep"] == sweep and ref_st[b]["sweep"] == sweep
                for k in ("iters", "active", "max_update", "diverged_index", "listed"):
                    assert st[b][k] == ref_st[b][k], (rep, b, k)
                assert st[b]["energy"] == pytest.approx(ref_st[b]["energy"], rel=1e-12)
        # other work on the same stream between replays
        compute_displacement_field(lives[0], cans[0], DisplacementField(ctx, lives[0].dims), params(max_iters=3))


@pytest.mark.parametrize("sweep", ["dense", "active"])
def test_batch_stop_rule_per_volume(ctx, batch, sweep):
    # A threshold that stops volume 0 at a known iteration: between its running minimum of
    # max |du| and the first clearly lower value.
    print(f"case: batch of 4, stop rule, {sweep} sweep", flush=True)
    live, can = batch[0]
    ms = [compute_displacement_field(live, can, DisplacementField(ctx, live.dims), params(max_iters=k))["max_update"]
          for k in range(1, 31)]
    kstar = next(k for k in range(2, 30) if ms[k] < min(ms[:k]) * 0.98)
    kp = params(sweep, max_iters=200, stop_max_update=0.5 * (ms[kstar] + min(ms[:kstar])))
    ref, ref_st = per_volume(ctx, batch, kp)
    assert ref_st[0]["iters"] == kstar + 1
    lives, cans = lists(batch)
    fields = fresh_fields(ctx, batch)
    for rep in range(2):
        for f in fields:
            f.initializeAllVoxels((0.0, 0.0, 0.0))
        st = batch_deform(lives, cans, fields, kp)
        print("iters", [s["iters"] for s in st], [s["iters"] for s in ref_st], flush=True)
        for b, f in enumerate(fields):
            assert st[b]["iters"] == ref_st[b]["iters"] and st[b]["max_update"] == ref_st[b]["max_update"]
            assert np.array_equal(f.download(), ref[b]), (rep, b)


@pytest.mark.parametrize("sweep", ["dense", "active"])
def test_batch_divergence_in_one_volume(ctx, batch, sweep):
    print(f"case: batch of 4, NaN planted in volume 1, {sweep} sweep", flush=True)
    kp = params(sweep, max_iters=6)
    inits = [np.zeros((3, int(np.prod(v[0].dims))), dtype=np.float32) for v in batch]
    phi, _ = batch[1][0].download()
    act = np.nonzero(np.abs(phi) < 0.5 * TRUNC)[0]
    k = int(act[len(act) // 3])
    inits[1][1, k] = np.nan
    lives, cans = lists(batch)
    fields = fresh_fields(ctx, batch, inits)
    with pytest.raises(KFError) as ei:
        batch_deform(lives, cans, fields, kp)
    assert ei.value.status == KF_E_DIVERGED
    # the reported voxel is the lowest non-finite update of the first diverging iteration,
    # as kf_deform reports it for that volume alone
    one = DisplacementField(ctx, batch[1][0].dims)
    one.upload(inits[1])
    with pytest.raises(KFError) as e1:
        compute_displacement_field(batch[1][0], batch[1][1], one, kp)
    idx = str(e1.value).split("diverged at voxel ")[1].split()[0].strip(")'\"")
    print("diverged", str(ei.value), "| single:", idx, flush=True)
    assert f"volume 1: deformation field diverged at voxel {idx}" in str(ei.value)
    good = [0, 2, 3]
    ref, _ = per_volume(ctx, [batch[b] for b in good], kp, [inits[b] for b in good])
    for r, b in zip(ref, good):
        assert np.array_equal(fields[b].download(), r), b


def test_autotune_and_retuned_shape_give_same_fields(ctx, batch):
    kp = params(max_iters=20)
    ref, _ = per_volume(ctx, batch, kp)
    lives, cans = lists(batch)
    fields = fresh_fields(ctx, batch)
    cfg = batch_autotune(lives, cans, fields, kp, force=True)
    print("tuned", cfg, flush=True)
    assert cfg["candidates"] >= 4 and cfg["us_per_iteration"] > 0
    assert cfg["voxels_per_thread"] in (1, 2, 4, 8) and cfg["blocks"] > 0
    # tuning launches write only the fields' scratch buffers
    assert all(not f.download().any() for f in fields)
    batch_deform(lives, cans, fields, kp, stats=False)
    for b, f in enumerate(fields):
        assert np.array_equal(f.download(), ref[b]), b


def test_every_tile_shape_gives_the_same_fields(ctx, batch):
    # float4 tiles (volumes with N a multiple of 4: 48^3 and 70x45x58 here) and scalar tiles,
    # persistent and one-block-per-tile grids: the per-voxel arithmetic is the same.
    kp = params(max_iters=12)
    ref, ref_st = per_volume(ctx, batch, kp)
    lives, cans = lists(batch)
    for vpt in (1, 2, 4, 8):
        for blocks in (0, 300):
            cfg = batch_autotune(lives, cans, fresh_fields(ctx, batch), kp, shape=(vpt, blocks))
            assert cfg["voxels_per_thread"] == vpt
            fields = fresh_fields(ctx, batch)
            st = batch_deform(lives, cans, fields, kp)
            for b, f in enumerate(fields):
                assert np.array_equal(f.download(), ref[b]), (vpt, blocks, b)
                assert st[b]["iters"] == ref_st[b]["iters"] and st[b]["active"] == ref_st[b]["active"]
    batch_autotune(lives, cans, fresh_fields(ctx, batch), kp, force=True)


def test_batch_argument_errors(ctx, batch):
    lives, cans = lists(batch)
    f = DisplacementField(ctx, lives[0].dims)
    with pytest.raises(KFError) as ei:
        batch_deform(lives[:2], cans[:2], [f, f], params(max_iters=2))
    assert ei.value.status == KF_E_INVALID_ARG
    g = DisplacementField(ctx, lives[1].dims)
    with pytest.raises(KFError) as ei:  # dims of field 0 do not match volume 1
        batch_deform(lives[:2], cans[:2], [g, f], params(max_iters=2))
    assert ei.value.status == KF_E_INVALID_ARG
    assert batch_deform([], [], [], params(max_iters=2)) == []


def test_active_batch_with_an_empty_volume(ctx, batch):
    """A volume with no surface (nothing integrated: no voxel can move) next to the others:
    its iterations still count, its field stays zero, and the rest match kf_deform."""
    print("case: batch of 4 + one 40^3 volume with nothing integrated, active sweep", flush=True)
    kw = dict(voxel_size=VOX, clip_distance=TRUNC, max_surface=TRUNC, dtype="bf16")
    e_live, e_can = SDF(ctx, (40, 40, 40), (0.0, 0.0, 0.5), **kw), SDF(ctx, (40, 40, 40), (0.0, 0.0, 0.5), **kw)
    vols = list(batch) + [(e_live, e_can)]
    kp = params("active", max_iters=12)
    ref, ref_st = per_volume(ctx, vols, kp)
    lives, cans = lists(vols)
    fields = fresh_fields(ctx, vols)
    st = batch_deform(lives, cans, fields, kp)
    assert st[-1]["listed"] == 0 and st[-1]["iters"] == 12 and st[-1]["energy"] == 0.0
    assert not fields[-1].download().any()
    for b, f in enumerate(fields):
        assert np.array_equal(f.download(), ref[b]), b
        assert st[b]["iters"] == ref_st[b]["iters"] and st[b]["listed"] == ref_st[b]["listed"], b


def test_active_batch_of_many_small_volumes(ctx):
    """Thirty volumes of 6^3 to 35^3 in one active-set batch (most one marking span each, grids
    whose voxel count is not a multiple of 64, empty ones in between, some starting from a
    nonzero field, which lists voxels away from the surface too): the listing of all volumes
    in four launches must give every volume the field and stats of its own kf_deform call."""
    edges = [6 + (7 * b) % 30 for b in range(30)]
    print(f"case: batch of 30 ellipsoid volumes, edges {min(edges)}..{max(edges)}, active sweep, 10 iterations",
          flush=True)
    vols = []
    kw = dict(voxel_size=VOX, clip_distance=TRUNC, max_surface=TRUNC, dtype="bf16")
    for b, (n, mn, d1, d2) in enumerate(scenes.batch_scene(30, seed=0x5EED, edges=edges)):
        if b % 7 == 3:
            vols.append((SDF(ctx, (n, n, n), mn, **kw), SDF(ctx, (n, n, n), mn, **kw)))  # nothing integrated
        else:
            vols.append(make_pair(ctx, (n, n, n), mn, VOX, TRUNC, d2, d1, "bf16"))
    inits = []
    for b, (live, _) in enumerate(vols):
        u = np.zeros((3, live.n), dtype=np.float32)
        if b % 5 == 1:
            n = live.dims[0]
            u3 = u.reshape(3, n, n, n)
            u3[:, n // 4:n // 2, 1:n // 3, n // 3:] = scenes.seeded_field((n, n, n), amplitude=0.3, seed=b).reshape(
                3, n, n, n)[:, n // 4:n // 2, 1:n // 3, n // 3:]
        inits.append(u)
    kp = params("active", max_iters=10)
    ref, ref_st = per_volume(ctx, vols, kp, inits)
    lives, cans = lists(vols)
    fields = fresh_fields(ctx, vols, inits)
    st = batch_deform(lives, cans, fields, kp)
    listed = [s["listed"] for s in st]
    print(f"listed per volume: {listed}", flush=True)
    assert sum(1 for x in listed if x == 0) >= 4 and sum(1 for x in listed if x > 0) >= 20
    for b, f in enumerate(fields):
        assert np.array_equal(f.download(), ref[b]), b
        for k in ("iters", "listed", "active", "max_update", "diverged_index"):
            assert st[b][k] == ref_st[b][k], (b, k, st[b][k], ref_st[b][k])
        assert st[b]["listed"] == listed_on_cpu(lives[b], inits[b], kp), b
