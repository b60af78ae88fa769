"""Precision sweep at 256^3: the torus pair (R 0.2 m, r 0.06 m, 2 cm twist) at 4 mm voxels,
SDFs stored as FP32, BF16 and FP16, displacements and accumulation in FP32, reductions in
FP64, 100 adjoint iterations.  Per storage type:

  * storage error of phi (metres) against the FP64 integration;
  * kernel error of u (voxels) against the FP64 oracle run on the same stored SDFs: the
    arithmetic of the fused kernel alone, held to 1e-3 voxel;
  * total error of u against the FP64 oracle run on the FP64 SDFs: storage and arithmetic.
"""
import numpy as np
import pytest

import oracle as O
from killingfusion_amd import SDF, DisplacementField, KillingParams, compute_displacement_field, scenes

from helpers import oparams, ogrid, oracle_integrate

pytestmark = pytest.mark.gpu

N = 256
VOXEL = 0.004
TRUNC = 3 * VOXEL
MIN3D = (-0.512, -0.512, 0.288)


def stats(e):
    e = np.asarray(e, dtype=np.float64)
    return e.max(), float(np.sqrt(np.mean(e * e)))


@pytest.fixture(scope="module")
def torus():
    print("case: 256^3 torus pair, FP64 oracle integrate + 100 adjoint iterations", flush=True)
    d_can, d_live = scenes.torus_pair()
    g = ogrid((N, N, N), MIN3D, VOXEL, TRUNC)
    lphi, lw = oracle_integrate(g, TRUNC, d_live)
    cphi, cw = oracle_integrate(g, TRUNC, d_can)
    kp = KillingParams.paper(VOXEL, 3.0, method="adjoint", max_iters=100)
    u64, st64 = O.adj_deform(lphi, cphi, np.zeros((3, g.n)), g, oparams(kp))
    print("oracle fp64:", st64, flush=True)
    return dict(g=g, lphi=lphi, lw=lw, cphi=cphi, cw=cw, kp=kp, u64=u64, st64=st64)


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
def test_precision_sweep_256(ctx, torus, dtype):
    print(f"case: precision sweep 256^3, {dtype} storage", flush=True)
    g, kp = torus["g"], torus["kp"]
    kw = dict(voxel_size=VOXEL, clip_distance=TRUNC, max_surface=TRUNC, dtype=dtype)
    live, can = SDF(ctx, (N, N, N), MIN3D, **kw), SDF(ctx, (N, N, N), MIN3D, **kw)
    live.upload(torus["lphi"].astype(np.float32), torus["lw"].astype(np.int32))
    can.upload(torus["cphi"].astype(np.float32), torus["cw"].astype(np.int32))
    qlive, _ = live.download()
    qcan, _ = can.download()
    touched = torus["lw"] > 0
    phi_err = np.abs(qlive[touched] - torus["lphi"][touched])
    phi_max, phi_rms = stats(phi_err)
    field = DisplacementField(ctx, (N, N, N))
    st = compute_displacement_field(live, can, field, kp)
    ug = field.download().astype(np.float64)
    uq, stq = O.adj_deform(qlive.astype(np.float64), qcan.astype(np.float64), np.zeros((3, g.n)), g, oparams(kp))
    band = np.any(uq != 0, axis=0) | np.any(torus["u64"] != 0, axis=0) | np.any(ug != 0, axis=0)
    k_max, k_rms = stats(np.abs(ug - uq).max(axis=0)[band])
    te = np.abs(ug - torus["u64"]).max(axis=0)[band]
    t_max, t_rms = stats(te)
    t_q = np.quantile(te, [0.99, 0.999])
    print(f"precision {dtype}: phi storage max {phi_max:.3e} m rms {phi_rms:.3e} m | "
          f"u vs oracle(same SDFs) max {k_max:.3e} rms {k_rms:.3e} vox | "
          f"u vs oracle(FP64 SDFs) max {t_max:.3e} rms {t_rms:.3e} p99 {t_q[0]:.3e} p99.9 {t_q[1]:.3e} "
          f"n>1e-3 {(te > 1e-3).sum()} vox | moved voxels {band.sum()} | "
          f"energy {st['energy']:.6e} vs {stq['energy']:.6e}", flush=True)
    assert st["iters"] == stq["iters"] == 100
    # storage: FP32 rounds metres; BF16 (8 significant bits) and FP16 (11) round
    # phi / (Max + eps) to nearest: half an ulp relative to |phi| (the integration leaves
    # |phi| up to tens of centimetres in front of the surface), FP16 subnormals aside
    rel = {"f32": 2.0 ** -24, "bf16": 2.0 ** -8, "f16": 2.0 ** -11}[dtype]
    assert (phi_err <= rel * 1.001 * np.abs(torus["lphi"][touched]) + 1e-9).all()
    # the fused kernel's own arithmetic (FP32 against FP64 on the same stored SDFs)
    assert k_max <= 1e-6
    assert st["energy"] == pytest.approx(stq["energy"], rel=1e-6)
    # storage and arithmetic together, just above the measured distribution (256^3 torus, 100
    # iterations, 47,950 moved voxels: BF16 max 5.8e-2, p99.9 7.5e-3, 1.4% above 1e-3; FP16 max
    # 7.7e-3, p99.9 1.0e-3, 0.1% above 1e-3; FP32 1e-8).  BF16 misses the 1e-3-voxel target: its
    # 8 significant bits of phi move the optimum, most at voxels whose band membership flips.
    bound = {"f32": (1e-6, 1e-6, 0.0), "bf16": (0.07, 1e-2, 0.02), "f16": (1e-2, 1.5e-3, 0.002)}[dtype]
    assert t_max <= bound[0] and t_q[1] <= bound[1]
    assert (te > 1e-3).sum() <= bound[2] * te.size
