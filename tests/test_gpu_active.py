"""Active-set sweep of the adjoint deform (kf_params.sweep) against the dense sweep, through the
C ABI.

A Jacobi step moves a voxel only when the live SDF at j + u_j lies in the band
(KillingFusion.cpp:272-274), and with u_j = 0 that reads phi_live[j]; so the voxels with
u_j = 0 and phi_live[j] outside the band never move, and the active-set sweep leaves them
out.  Both sweeps run the same per-voxel arithmetic, so the fields (compared as bit
patterns) and the iteration statistics must agree exactly.  The final energy has the same
per-voxel terms summed in another order: each sweep's is checked against the exactly rounded
sum of the dense forward pass's terms.  The dense sweep itself is held to the FP64 oracle in
test_gpu_parity.py.
"""
import dataclasses
import math

import numpy as np
import pytest

from killingfusion_amd import DisplacementField, KillingParams, compute_displacement_field, energy_forward, scenes
from killingfusion_amd._lib import KF_E_DIVERGED, KFError

from helpers import C1, listed_on_cpu, make_pair, sphere_frames

pytestmark = pytest.mark.gpu

VOX = 0.006
TRUNC = 3 * VOX


def run(ctx, live, can, kp, sweep, u0=None):
    """Deforms with the given sweep; also checks the reported final energy against the exactly
    rounded sum (math.fsum) of the dense forward pass's per-voxel terms on the final field,
    within the recursive-summation bound (n - 1) 2^-53 sum |e_i| (so a missing or changed
    term fails it, a different summation order does not)."""
    f = DisplacementField(ctx, live.dims)
    if u0 is not None:
        f.upload(u0)
    try:
        st = compute_displacement_field(live, can, f, dataclasses.replace(kp, sweep=sweep))
        _, ev = energy_forward(live, can, f, kp, per_voxel=True)
        ev = ev.astype(np.float64)
        exact = math.fsum(ev)
        bound = max(int(np.count_nonzero(ev)) - 1, 0) * 2.0 ** -53 * float(np.abs(ev).sum())
        assert abs(st["energy"] - exact) <= bound, (sweep, st["energy"], exact, bound)
        return f.download(), st
    finally:
        f.close()


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def assert_same(a, b, tag):
    (ua, sa), (ub, sb) = a, b
    print(f"{tag}: iters {sb['iters']} listed {sb['listed']} active {sb['active']} max {sb['max_update']:.3e}",
          flush=True)
    assert sa["sweep"] == "dense" and sb["sweep"] == "active"
    assert same_bits(ua, ub), f"{tag}: fields differ"
    for k in ("iters", "max_update", "active", "diverged_index"):
        assert sa[k] == sb[k], (tag, k, sa[k], sb[k])


@pytest.fixture(scope="module")
def sphere(ctx):
    d1, d2 = sphere_frames()
    out = {}
    for dt in ("f32", "bf16", "f16"):
        out[dt] = make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype=dt)
    return out


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_active_equals_dense_sphere(ctx, sphere, dtype):
    print(f"case: 64^3 sphere pair, {dtype}, 60 adjoint iterations, dense vs active", flush=True)
    live, can = sphere[dtype]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=60)
    a = run(ctx, live, can, kp, "dense")
    b = run(ctx, live, can, kp, "active")
    assert_same(a, b, f"sphere {dtype}")
    assert b[1]["listed"] == listed_on_cpu(live, None, kp)
    assert 0 < b[1]["active"] <= b[1]["listed"] < live.n // 20


@pytest.mark.parametrize("terms", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1)])
def test_active_equals_dense_energy_terms(ctx, sphere, terms):
    """Each energy term alone and in pairs (EnergyTypeUsed, config.cpp:23)."""
    d, s_, k = terms
    print(f"case: 64^3 sphere, f32, data {d} level set {s_} Killing {k}, 40 iterations", flush=True)
    live, can = sphere["f32"]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=40, use_data=bool(d),
                             use_levelset=bool(s_), use_killing=bool(k))
    assert_same(run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active"), f"terms {terms}")


@pytest.mark.parametrize("dims", [(61, 45, 58), (70, 45, 58), (48, 48, 48)])
def test_active_equals_dense_uneven(ctx, dims):
    """Odd and even shapes: N % 4 != 0 runs the dense one-voxel-per-thread kernel."""
    print(f"case: ellipsoid pair {dims}, 6 mm, bf16, 50 iterations, dense vs active", flush=True)
    ext = [d * VOX for d in dims]
    d1, d2, _ = scenes.ellipsoid_pair(0x5EED, 3, max_axis=0.4 * min(ext))
    live, can = make_pair(ctx, dims, (-ext[0] / 2, -ext[1] / 2, 0.8 - ext[2] / 2), VOX, TRUNC, d2, d1, "bf16")
    kp = KillingParams.paper(VOX, 3.0, method="adjoint", max_iters=50)
    assert_same(run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active"), f"ellipsoid {dims}")


@pytest.mark.parametrize("dims", [(61, 45, 58), (63, 47, 29)])
def test_active_with_initial_field_uneven(ctx, dims):
    """A nonzero starting field in a sub-box on grids whose voxel count is not a multiple of 4
    (the marking pass's one-voxel-per-lane path): the sweeps agree bit for bit and the list
    holds exactly the voxels with a nonzero u bit pattern or an in-band live value."""
    print(f"case: ellipsoid pair {dims}, 6 mm, bf16, start field in a sub-box, 30 iterations", flush=True)
    ext = [d * VOX for d in dims]
    d1, d2, _ = scenes.ellipsoid_pair(0x5EED, 5, max_axis=0.4 * min(ext))
    live, can = make_pair(ctx, dims, (-ext[0] / 2, -ext[1] / 2, 0.8 - ext[2] / 2), VOX, TRUNC, d2, d1, "bf16")
    kp = KillingParams.paper(VOX, 3.0, method="adjoint", max_iters=30)
    u0 = scenes.seeded_field(dims, amplitude=0.4, seed=11).reshape(3, dims[2], dims[1], dims[0])
    part = np.zeros_like(u0)
    part[:, 3:17, 20:41, 7:33] = u0[:, 3:17, 20:41, 7:33]
    part[1, dims[2] - 1, dims[1] - 1, dims[0] - 1] = -0.0  # the last voxel: listed, never moves
    part = part.reshape(3, -1)
    b = run(ctx, live, can, kp, "active", part)
    assert_same(run(ctx, live, can, kp, "dense", part), b, f"sub-box field {dims}")
    assert b[1]["listed"] == listed_on_cpu(live, part, kp)


def test_active_with_initial_field(ctx, sphere):
    """A nonzero starting field in a sub-box (voxels far from the surface move too), then a field
    nonzero everywhere (every voxel listed; the automatic choice falls back to dense)."""
    live, can = sphere["f32"]
    dims = C1["dims"]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=30)
    u0 = scenes.seeded_field(dims, amplitude=0.4, seed=7).reshape(3, dims[2], dims[1], dims[0])
    part = np.zeros_like(u0)
    part[:, 5:20, 30:50, 10:40] = u0[:, 5:20, 30:50, 10:40]
    part[0, 0, 0, 0] = -0.0  # listed (not +0), though it never moves
    part = part.reshape(3, -1)
    print("case: 64^3 sphere, f32, initial field in a sub-box, 30 iterations", flush=True)
    b = run(ctx, live, can, kp, "active", part)
    assert_same(run(ctx, live, can, kp, "dense", part), b, "sub-box field")
    assert b[1]["listed"] == listed_on_cpu(live, part, kp)
    full = u0.reshape(3, -1)
    print("case: 64^3 sphere, f32, initial field nonzero everywhere, 30 iterations", flush=True)
    b = run(ctx, live, can, kp, "active", full)
    assert_same(run(ctx, live, can, kp, "dense", full), b, "full field")
    assert b[1]["listed"] == live.n
    _, auto = run(ctx, live, can, kp, "auto", full)
    assert auto["sweep"] == "dense" and auto["listed"] == live.n
    _, auto = run(ctx, live, can, kp, "auto")
    assert auto["sweep"] == "active"


def test_active_stop_rule(ctx, sphere):
    live, can = sphere["bf16"]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=400, stop_max_update=2e-2)
    print("case: 64^3 sphere, bf16, stop rule 2e-2 voxel, dense vs active", flush=True)
    a = run(ctx, live, can, kp, "dense")
    b = run(ctx, live, can, kp, "active")
    assert_same(a, b, "stop rule")
    assert 2 <= b[1]["iters"] < 400 and b[1]["max_update"] < 2e-2


def test_active_divergence(ctx, sphere):
    """A NaN planted in the starting field at an in-band voxel: both sweeps stop with the same
    lowest non-finite voxel."""
    live, can = sphere["f32"]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=8)
    phi, _ = live.download()
    act = np.nonzero(np.abs(phi) < 0.5 * C1["trunc"])[0]
    u0 = np.zeros((3, live.n), dtype=np.float32)
    u0[2, int(act[len(act) // 2])] = np.nan
    msgs = []
    for sweep in ("dense", "active"):
        print(f"case: 64^3 sphere, f32, NaN in the starting field, {sweep}", flush=True)
        f = DisplacementField(ctx, C1["dims"])
        f.upload(u0)
        with pytest.raises(KFError) as e:
            compute_displacement_field(live, can, f, dataclasses.replace(kp, sweep=sweep))
        assert e.value.status == KF_E_DIVERGED
        msgs.append(str(e.value))
        f.close()
    print(msgs, flush=True)
    assert msgs[0] == msgs[1]


def test_active_equals_dense_256(ctx):
    """The benchmark's workload: 256^3 torus pair, 4 mm voxels, BF16, 100 iterations."""
    print("case: 256^3 torus pair, bf16, 100 iterations, dense vs active", flush=True)
    d_can, d_live = scenes.torus_pair()
    dims = (256, 256, 256)
    live, can = make_pair(ctx, dims, (-0.512, -0.512, 0.288), 0.004, 0.012, d_live, d_can, "bf16")
    kp = KillingParams.paper(0.004, 3.0, method="adjoint", max_iters=100)
    a = run(ctx, live, can, kp, "dense")
    b = run(ctx, live, can, kp, "active")
    assert_same(a, b, "torus 256^3")
    assert b[1]["listed"] == listed_on_cpu(live, None, kp)
    live.close()
    can.close()


def test_active_loop_replayed_as_a_graph(loop_capture, sphere):
    """With loop capture on and without a stop rule, the iteration loop of an active-set call
    runs directly the first time its launch shape is seen, is captured as a hipGraph the second
    time and replayed after that:
    every call gives the dense sweep's field and statistics bit for bit, with other work on the
    stream in between (a call of another shape, which the cache must not mistake for this one)."""
    ctx = loop_capture
    live, can = sphere["bf16"]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=30)
    ref = run(ctx, live, can, kp, "dense")
    l32, c32 = sphere["f32"]
    for call in range(5):
        print(f"case: 64^3 sphere, bf16, 30 iterations, active call {call} (direct, captured, replayed)", flush=True)
        assert_same(ref, run(ctx, live, can, kp, "active"), f"active call {call}")
        if call == 2:  # a different shape in between, then this one again
            run(ctx, l32, c32, dataclasses.replace(kp, max_iters=7), "active")


def test_list_longer_than_predicted(ctx, sphere):
    """A call sizes its list launches from the list length of the previous call on a volume of
    the same size, without reading the length back first.  A start field that is nonzero over
    a sub-box lists far more voxels than the call before it: the list kernels then do nothing
    and the call runs again with the length read back.  Every call still gives the dense
    sweep's field and statistics bit for bit, and so does a short list right after the long
    one (launched over more slots than it needs)."""
    live, can = sphere["bf16"]
    dims = C1["dims"]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint", max_iters=12)
    u0 = scenes.seeded_field(dims, amplitude=0.3, seed=4).reshape(3, dims[2], dims[1], dims[0])
    box = np.zeros_like(u0)
    box[:, 10:50, 5:60, 5:60] = u0[:, 10:50, 5:60, 5:60]
    box = box.reshape(3, -1)
    print("case: 64^3 sphere, bf16, short list, then a long one, then short again", flush=True)
    short_d = run(ctx, live, can, kp, "dense")
    short_a = run(ctx, live, can, kp, "active")  # sets the prediction
    assert_same(short_d, short_a, "short list")
    long_a = run(ctx, live, can, kp, "active", box)
    assert long_a[1]["listed"] > 2 * short_a[1]["listed"] + 4096
    assert long_a[1]["listed"] == listed_on_cpu(live, box, kp)
    assert_same(run(ctx, live, can, kp, "dense", box), long_a, "long list after a short one")
    again = run(ctx, live, can, kp, "active")
    assert_same(short_d, again, "short list after a long one")


@pytest.mark.parametrize("dtype", ["bf16", "f32"])
def test_dense_calls_split_on_one_field(ctx, sphere, dtype):
    """The dense step stores a voxel group only where its destination buffer can differ from
    the new value (the groups the stencil reached in the step before), and stores every group
    in the first step of a call, whatever the field's other buffer holds.  Twelve iterations
    as one call, as three calls on the same field, and with an active-set call (which leaves
    the other buffer stale) between dense ones, give the same field bit for bit."""
    print(f"case: dense deform split over calls on one field, {dtype}", flush=True)
    live, can = sphere[dtype]
    kp = KillingParams.paper(C1["voxel"], 3.0, method="adjoint")
    whole, _ = run(ctx, live, can, dataclasses.replace(kp, max_iters=12), "dense")
    f = DisplacementField(ctx, live.dims)
    try:
        plans = ((3, "dense"), (4, "dense"), (5, "dense"), (3, "dense"), (4, "active"), (5, "dense"))
        for start in (0, 3):
            f.initializeAllVoxels((0.0, 0.0, 0.0))
            for iters, sweep in plans[start:start + 3]:
                compute_displacement_field(live, can, f, dataclasses.replace(kp, max_iters=iters, sweep=sweep))
            assert same_bits(f.download(), whole), f"split {plans[start:start + 3]}: fields differ"
    finally:
        f.close()
