"""Active-set sweep of the reference-formula optimiser (kf_params.sweep with the reference
method) against its dense sweep, through the C ABI.

The reference's Jacobi step (KillingFusion.cpp:248-351) updates voxel j only when the warped
live SDF at j + u_j is in the band (:272-274), and with u_j = 0 and every live value finite
that value is phi_live[j]; a skipped voxel takes u_j plus the +0 of its iteration's fresh delta
field (:254-256, 345-348), so it keeps u_j except that a -0 component becomes +0.  So the voxels with
u_j = +0 and phi_live[j] outside the band never move, and the active-set sweep visits only the
others, in the field's own two buffers.  (With a non-finite live value the warped value next
to it is NaN, which passes the band test; the sweep then steps densely.)  Both sweeps run the
same per-voxel arithmetic, so the fields (compared as bit patterns) and the statistics must
agree exactly; the dense sweep itself is held to the FP64 oracle in test_gpu_parity.py and
test_gpu_reference.py.
"""
import dataclasses

import numpy as np
import pytest

from killingfusion_amd import DisplacementField, KillingParams, compute_displacement_field, scenes
from killingfusion_amd._lib import KF_E_DIVERGED, KFError

from helpers import C1, listed_on_cpu, make_pair, sphere_frames

pytestmark = pytest.mark.gpu


def run(ctx, live, can, kp, sweep, u0=None):
    f = DisplacementField(ctx, live.dims)
    if u0 is not None:
        f.upload(u0)
    try:
        st = compute_displacement_field(live, can, f, dataclasses.replace(kp, sweep=sweep))
        return f.download(), st
    finally:
        f.close()


def same(a, b, tag):
    (ua, sa), (ub, sb) = a, b
    print(f"{tag}: iters {sb['iters']} listed {sb['listed']} active {sb['active']} max {sb['max_update']:.3e}",
          flush=True)
    assert sa["sweep"] == "dense" and sb["sweep"] == "active"
    assert np.array_equal(ua.view(np.uint32), ub.view(np.uint32)), f"{tag}: fields differ"
    for k in ("iters", "max_update", "active", "diverged_index"):
        assert sa[k] == sb[k], (tag, k, sa[k], sb[k])


@pytest.fixture(scope="module")
def sphere(ctx):
    d1, d2 = sphere_frames()
    return {dt: make_pair(ctx, C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"], d2, d1, dtype=dt)
            for dt in ("f32", "bf16")}


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("params", ["config", "paper"])
def test_reference_active_equals_dense_sphere(ctx, sphere, dtype, params):
    print(f"case: reference method, 64^3 sphere, {dtype}, {params} parameters, dense vs active", flush=True)
    live, can = sphere[dtype]
    v, t = C1["voxel"], C1["trunc"]
    kp = (KillingParams(voxel_size=v, max_surface=t, clip_distance=t, max_iters=40) if params == "config"
          else KillingParams.paper(v, 3.0, max_iters=40))
    a, b = run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active")
    same(a, b, f"sphere {dtype} {params}")
    assert b[1]["listed"] == listed_on_cpu(live, None, kp) and 0 < b[1]["active"] <= b[1]["listed"]
    _, auto = run(ctx, live, can, kp, "auto")
    assert auto["sweep"] == "active"


@pytest.mark.parametrize("dims", [(61, 45, 58), (48, 48, 48)])
def test_reference_active_equals_dense_uneven_with_start_field(ctx, dims):
    print(f"case: reference method, ellipsoid pair {dims}, bf16, start field in a sub-box", flush=True)
    vox = 0.006
    ext = [d * vox for d in dims]
    d1, d2, _ = scenes.ellipsoid_pair(0x5EED, 3, max_axis=0.4 * min(ext))
    live, can = make_pair(ctx, dims, (-ext[0] / 2, -ext[1] / 2, 0.8 - ext[2] / 2), vox, 3 * vox, d2, d1, "bf16")
    kp = KillingParams.paper(vox, 3.0, max_iters=25)
    u0 = scenes.seeded_field(dims, amplitude=0.4, seed=9).reshape(3, dims[2], dims[1], dims[0])
    part = np.zeros_like(u0)
    part[:, 4:15, 10:30, 5:35] = u0[:, 4:15, 10:30, 5:35]
    part[2, 0, 0, 1] = -0.0
    part = part.reshape(3, -1)
    for start in (None, part):
        a, b = run(ctx, live, can, kp, "dense", start), run(ctx, live, can, kp, "active", start)
        same(a, b, f"ellipsoid {dims} start {'zero' if start is None else 'sub-box'}")
        assert b[1]["listed"] == listed_on_cpu(live, start, kp)


def test_reference_active_stop_rule_and_divergence(ctx, sphere):
    live, can = sphere["f32"]
    # a threshold just above the largest update of iteration 10 stops the call by then
    m10 = run(ctx, live, can, KillingParams.paper(C1["voxel"], 3.0, max_iters=10), "dense")[1]["max_update"]
    kp = KillingParams.paper(C1["voxel"], 3.0, max_iters=300, stop_max_update=m10 * 1.0001)
    print(f"case: reference method, 64^3 sphere, stop rule {m10 * 1.0001:.3e} voxel, dense vs active", flush=True)
    a, b = run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active")
    same(a, b, "stop rule")
    assert 1 <= b[1]["iters"] <= 10
    phi, _ = live.download()
    act = np.nonzero(np.abs(phi) < 0.5 * C1["trunc"])[0]
    u0 = np.zeros((3, live.n), dtype=np.float32)
    u0[1, int(act[len(act) // 3])] = np.nan
    msgs = []
    for sweep in ("dense", "active"):
        print(f"case: reference method, NaN in the start field, {sweep}", flush=True)
        f = DisplacementField(ctx, C1["dims"])
        f.upload(u0)
        with pytest.raises(KFError) as e:
            compute_displacement_field(live, can, f, KillingParams.paper(C1["voxel"], 3.0, max_iters=6, sweep=sweep))
        assert e.value.status == KF_E_DIVERGED
        msgs.append(str(e.value))
        f.close()
    assert msgs[0] == msgs[1]


def test_reference_sweep_with_a_non_finite_live_value(ctx, sphere):
    """With a non-finite live value the warped value of a voxel next to it is NaN even where
    u = +0, and NaN passes the band test (KillingFusion.cpp:272-274): such a voxel moves though
    phi_live[j] is outside the band, so the list would miss it.  The active sweep then steps
    densely, and both requests give the dense result (here the divergence at that voxel)."""
    live, can = sphere["f32"]
    kp = KillingParams(voxel_size=C1["voxel"], max_surface=C1["trunc"], clip_distance=C1["trunc"], max_iters=5)
    # a finite call first, so that the calls below size their list launches from its list
    assert run(ctx, live, can, kp, "active")[1]["sweep"] == "active"
    phi, w = live.download()
    k = 64 * 64 * 40 + 64 * 20 + 33
    p3 = phi.copy()
    p3[k] = np.inf
    live.upload(p3, w)
    try:
        msgs = []
        for sweep in ("dense", "active"):
            print(f"case: reference method, +inf live value at voxel {k}, {sweep}", flush=True)
            f = DisplacementField(ctx, C1["dims"])
            with pytest.raises(KFError) as e:
                compute_displacement_field(live, can, f, dataclasses.replace(kp, sweep=sweep))
            assert e.value.status == KF_E_DIVERGED
            msgs.append(str(e.value))
            f.close()
        assert msgs[0] == msgs[1]
    finally:
        live.upload(phi, w)


def test_reference_active_equals_dense_256(ctx):
    """The benchmark's workload: the 256^3 torus pair, FP32, paper parameters, 20 iterations."""
    print("case: reference method, 256^3 torus pair, f32, 20 iterations, dense vs active", flush=True)
    d_can, d_live = scenes.torus_pair()
    live, can = make_pair(ctx, (256, 256, 256), (-0.512, -0.512, 0.288), 0.004, 0.012, d_live, d_can, "f32")
    kp = KillingParams.paper(0.004, 3.0, max_iters=20)
    same(run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active"), "torus 256^3")
    live.close()
    can.close()


def test_reference_list_longer_than_predicted(ctx, sphere):
    """The reference method's active sweep sizes its list launches from the previous call's
    list as well: a start field nonzero over a sub-box lists far more voxels than the call
    before, the list kernels do nothing and the call runs again with the length read back;
    the field is the dense sweep's bit for bit."""
    live, can = sphere["bf16"]
    dims = C1["dims"]
    kp = KillingParams.paper(C1["voxel"], 3.0, max_iters=10)
    u0 = scenes.seeded_field(dims, amplitude=0.3, seed=5).reshape(3, dims[2], dims[1], dims[0])
    box = np.zeros_like(u0)
    box[:, 8:52, 4:60, 6:58] = u0[:, 8:52, 4:60, 6:58]
    box = box.reshape(3, -1)
    print("case: reference method, 64^3 sphere, bf16, short list then a long one", flush=True)
    short = run(ctx, live, can, kp, "active")
    long_a = run(ctx, live, can, kp, "active", box)
    assert long_a[1]["listed"] > 2 * short[1]["listed"] + 4096
    same(run(ctx, live, can, kp, "dense", box), long_a, "reference, long list after a short one")
    same(run(ctx, live, can, kp, "dense"), run(ctx, live, can, kp, "active"), "reference, short list again")


def test_skipped_voxels_turn_minus_zero_into_plus_zero(ctx, sphere):
    """The reference adds each iteration's delta field, created as zeros (+0), to u
    (KillingFusion.cpp:254-256, 345-348); a voxel outside the band gets no update, so its -0
    components become +0 after one iteration, as in the oracle.  Both sweeps, bit for bit."""
    import oracle as O
    from helpers import ogrid, oparams
    live, can = sphere["f32"]
    phi, _ = live.download()
    cphi, _ = can.download()
    far = np.nonzero(np.abs(phi) > 2 * C1["trunc"])[0][::97][:64]
    u0 = np.zeros((3, live.n), dtype=np.float32)
    u0[:, far] = -0.0
    kp = KillingParams.paper(C1["voxel"], 3.0, max_iters=1)
    g = ogrid(C1["dims"], C1["min3d"], C1["voxel"], C1["trunc"])
    uo, _ = O.deform_jacobi(phi.astype(np.float64), cphi.astype(np.float64), u0.astype(np.float64), g, oparams(kp))
    assert not np.signbit(uo[:, far]).any()
    for sweep in ("dense", "active"):
        print(f"case: reference method, -0 start components outside the band, {sweep}", flush=True)
        ug, st = run(ctx, live, can, kp, sweep, u0)
        assert st["sweep"] == sweep
        assert (ug[:, far].view(np.uint32) == 0).all(), sweep
