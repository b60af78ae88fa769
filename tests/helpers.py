"""Shared workload setup for the parity tests: the same synthetic inputs go to libkf (GPU)
and to the FP64 oracle (CPU)."""
from __future__ import annotations

import functools

import numpy as np

import oracle as O
from killingfusion_amd import scenes

# Sphere pair at 64^3, 8 mm voxels, truncation 3 voxels (the CPU-runnable workload).
C1 = dict(dims=(64, 64, 64), min3d=(-0.256, -0.256, 0.544), voxel=0.008, trunc=0.024)


@functools.lru_cache(maxsize=None)
def sphere_frames():
    return scenes.sphere_pair()


def ogrid(dims, min3d, voxel, trunc, eps=1e-5):
    return O.OGrid(dims=tuple(dims), min_corner=tuple(min3d), voxel=voxel, oob=trunc + eps)


def oparams(kp) -> O.OParams:
    """Oracle params from a KillingParams."""
    return O.OParams(voxel_size=kp.voxel_size, max_surface=kp.max_surface, clip=kp.clip_distance,
                     epsilon=kp.epsilon, alpha=kp.alpha, omega_k=kp.omega_killing, gamma=kp.gamma_killing,
                     omega_s=kp.omega_levelset, h=kp.delta_size, killing_clamp=kp.killing_clamp,
                     use_data=kp.use_data, use_levelset=kp.use_levelset, use_killing=kp.use_killing,
                     max_iters=kp.max_iters, stop_max_update=kp.stop_max_update,
                     use_trust=getattr(kp, "use_trust", False))


def oracle_integrate(g, trunc, depth):
    phi, w = O.sdf_init(g)
    O.integrate(phi, w, g, trunc, depth, scenes.K, scenes.C2W, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    return phi, w


def bf16_roundtrip(phi_m: np.ndarray, scale: float) -> np.ndarray:
    """What a BF16 volume stores and returns: RNE(phi / scale) in bfloat16, times scale (FP32)."""
    s = np.float32(scale)
    q = (np.asarray(phi_m, dtype=np.float32) / s).astype(np.float32)
    b = q.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return (r.view(np.float32) * s).astype(np.float32)


def f16_roundtrip(phi_m: np.ndarray, scale: float) -> np.ndarray:
    s = np.float32(scale)
    q = (np.asarray(phi_m, dtype=np.float32) / s).astype(np.float32)
    return (q.astype(np.float16).astype(np.float32) * s).astype(np.float32)


def make_pair(ctx, dims, min3d, voxel, trunc, d_live, d_can, dtype="f32"):
    """GPU live/canonical SDFs integrated from two depth frames."""
    from killingfusion_amd import SDF
    kw = dict(voxel_size=voxel, clip_distance=trunc, max_surface=trunc, dtype=dtype)
    live = SDF(ctx, dims, min3d, **kw)
    can = SDF(ctx, dims, min3d, **kw)
    live.integrateDepthFrame(d_live, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    can.integrateDepthFrame(d_can, scenes.C2W, scenes.K, scenes.MIN_DEPTH, scenes.MAX_DEPTH)
    return live, can


def listed_on_cpu(live, u0, kp):
    """The voxels an active-set sweep lists: some component of u_j with a nonzero bit pattern
    (-0 included) or phi_live[j] in the band -clip <= phi <= Max - eps (KillingFusion.cpp:273),
    compared in FP32 as the kernels compare them."""
    phi, _ = live.download()
    hi = np.float32(kp.max_surface - kp.epsilon)
    lo = np.float32(-kp.clip_distance)
    band = ~((phi > hi) | (phi < lo))
    if u0 is None:
        return int(band.sum())
    moved = (np.asarray(u0, np.float32).view(np.uint32) != 0).any(axis=0)
    return int((band | moved).sum())
