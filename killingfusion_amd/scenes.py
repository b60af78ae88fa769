"""Synthetic depth frames and seeded fields for the hot-path workloads.

The reference reads the Slavcheva deformable dataset from PNG files
(KillingFusion/src/DatasetReader.cpp:50-67, 122-134: 640x480 depth in mm -> metres).
These generators render analytic scenes instead, as float32 depth in metres, with a
pinhole camera K = [[525,0,319.5],[0,525,239.5],[0,0,1]] and camera-to-world pose I
(KillingFusion.cpp:225-229 passes Identity).  Pixel (col, row) looks along
((col-cx)/fx, (row-cy)/fy, 1); a ray that hits nothing stores 0 (invalid, rejected by
the min-depth test of SDF.cpp:178).

All randomness comes from a project-written splitmix64 stream, so every workload is
bit-reproducible across machines (numpy float64 math, cast to float32 at the end).
"""
from __future__ import annotations

import math

import numpy as np

WIDTH, HEIGHT = 640, 480
FX = FY = 525.0
CX, CY = 319.5, 239.5
K = np.array([[FX, 0.0, CX], [0.0, FY, CY], [0.0, 0.0, 1.0]], dtype=np.float64)
C2W = np.eye(4, dtype=np.float64)
MIN_DEPTH, MAX_DEPTH = 0.1, 3.0

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """The splitmix64 sequence (Steele, Lea, Flood 2014): element i is the output after
    i+1+offset increments of the state seed, as uint64."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform [0, 1) doubles from the top 53 bits of splitmix64."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller on pairs of uniforms."""
    m = (count + 1) // 2
    u = uniform(seed, 2 * m, 2 * offset)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2 * math.pi * u2)
    out[1::2] = r * np.sin(2 * math.pi * u2)
    return out[:count]


def _rays():
    col, row = np.meshgrid(np.arange(WIDTH, dtype=np.float64), np.arange(HEIGHT, dtype=np.float64))
    dx = (col - CX) / FX
    dy = (row - CY) / FY
    return dx.ravel(), dy.ravel()


def _march(sdf, z0: float, z1: float, lipschitz: float = 1.0, iters: int = 400, tol: float = 1e-7,
           bound=None):
    """Sphere-trace f (negative inside) along every pixel ray parametrised by camera z.
    `lipschitz` bounds |grad f| along the ray per unit z (>= 1 for a true SDF; the ray
    direction has length >= 1 per unit z).  `bound` = (centre, radius), a sphere holding the
    whole surface: rays that miss it are misses, the others start where they enter it.
    Returns float32 depth (0 where no hit)."""
    dx, dy = _rays()
    scale = np.sqrt(dx * dx + dy * dy + 1.0)  # ray length per unit z
    z = np.full(dx.shape, z0)
    alive = np.ones(dx.shape, dtype=bool)
    if bound is not None:
        # |(dx, dy, 1) z - c|^2 = R^2 as a quadratic in z: a z^2 - 2 b z + cc = 0
        c = np.asarray(bound[0], dtype=np.float64)
        a = scale * scale
        b = dx * c[0] + dy * c[1] + c[2]
        cc = float(c @ c) - float(bound[1]) ** 2
        disc = b * b - a * cc
        alive = disc >= 0.0
        z = np.where(alive, np.maximum(z0, (b - np.sqrt(np.maximum(disc, 0.0))) / a), z0)
    hit = np.zeros(dx.shape, dtype=bool)
    for _ in range(iters):
        idx = np.nonzero(alive)[0]
        if idx.size == 0:
            break
        zi = z[idx]
        p = np.stack([dx[idx] * zi, dy[idx] * zi, zi], axis=1)
        f = sdf(p)
        conv = f < tol
        hit[idx[conv]] = True
        alive[idx[conv]] = False
        step = f / (scale[idx] * lipschitz)
        znew = zi + np.maximum(step, 0.0)
        z[idx] = np.where(conv, zi, znew)
        far = (~conv) & (znew > z1)
        alive[idx[far]] = False
    # Bisection refinement between the last outside point and just inside.
    hi = np.nonzero(hit)[0]
    if hi.size:
        lo_z = z[hi] - 1e-4
        hi_z = z[hi] + 1e-4
        p_lo = np.stack([dx[hi] * lo_z, dy[hi] * lo_z, lo_z], axis=1)
        ok = sdf(p_lo) > 0
        lo_z = np.where(ok, lo_z, z[hi] - 5e-3)
        for _ in range(40):
            mid = 0.5 * (lo_z + hi_z)
            pm = np.stack([dx[hi] * mid, dy[hi] * mid, mid], axis=1)
            inside = sdf(pm) <= 0
            hi_z = np.where(inside, mid, hi_z)
            lo_z = np.where(inside, lo_z, mid)
        z[hi] = 0.5 * (lo_z + hi_z)
    depth = np.where(hit, z, 0.0)
    return depth.reshape(HEIGHT, WIDTH).astype(np.float32)


# ---------------------------------------------------------------------------
# Shapes (implicit functions, metres, camera frame)
# ---------------------------------------------------------------------------

def sphere_sdf(center, radius):
    c = np.asarray(center, dtype=np.float64)

    def f(p):
        return np.linalg.norm(p - c, axis=1) - radius
    return f


def bulged_sphere_fn(center, radius, amp, m_theta=3, m_phi=2):
    """|p - c| - (r + amp sin(m_theta*theta) cos(m_phi*phi)), theta polar from +z."""
    c = np.asarray(center, dtype=np.float64)

    def f(p):
        d = p - c
        rr = np.linalg.norm(d, axis=1)
        theta = np.arccos(np.clip(d[:, 2] / np.maximum(rr, 1e-12), -1, 1))
        ph = np.arctan2(d[:, 1], d[:, 0])
        return rr - (radius + amp * np.sin(m_theta * theta) * np.cos(m_phi * ph))
    return f


def _rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], dtype=np.float64)


def torus_fn(center, R, r, tilt_deg=30.0, twist=0.0):
    """Torus with axis along local z, tilted by tilt_deg about x.  `twist` (metres) rotates
    each local point about the local x axis by angle twist*x/(R*(R+r)): points at the rim
    (x = +-R) move by about `twist`."""
    c = np.asarray(center, dtype=np.float64)
    Rm = _rot_x(math.radians(tilt_deg))

    def f(p):
        q = (p - c) @ Rm  # world -> local (Rm orthonormal: local = Rm^T (p - c))
        if twist != 0.0:
            ang = -twist * q[:, 0] / (R * (R + r))
            ca, sa = np.cos(ang), np.sin(ang)
            y = ca * q[:, 1] - sa * q[:, 2]
            zz = sa * q[:, 1] + ca * q[:, 2]
            q = np.stack([q[:, 0], y, zz], axis=1)
        a = np.sqrt(q[:, 0] ** 2 + q[:, 1] ** 2) - R
        return np.sqrt(a * a + q[:, 2] ** 2) - r
    return f


def ellipsoid_fn(center, axes, warp_terms=()):
    """Deformed ellipsoid: f(p) = (|A^-1 (p - c - w(p))| - 1) * min(axes), w = sum of
    amplitude * sin(k . p + phase) along a direction."""
    c = np.asarray(center, dtype=np.float64)
    ax = np.asarray(axes, dtype=np.float64)

    def f(p):
        q = p - c
        for amp, kvec, phase, dirv in warp_terms:
            q = q - amp * np.sin(q @ kvec + phase)[:, None] * dirv[None, :]
        return (np.linalg.norm(q / ax, axis=1) - 1.0) * ax.min()
    return f


def capsule_sdf(a, b, radius):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ab = b - a
    L2 = float(ab @ ab)

    def f(p):
        t = np.clip((p - a) @ ab / L2, 0.0, 1.0)
        closest = a[None, :] + t[:, None] * ab[None, :]
        return np.linalg.norm(p - closest, axis=1) - radius
    return f


def union(*fs):
    def f(p):
        return np.min(np.stack([g(p) for g in fs], axis=0), axis=0)
    return f


# ---------------------------------------------------------------------------
# Workload scenes
# ---------------------------------------------------------------------------

def sphere_pair():
    """Sphere r = 0.15 m at (0, 0, 0.8); frame 2 adds a radial bulge 0.01 sin(3 theta) cos(2 phi)."""
    c = (0.0, 0.0, 0.8)
    d1 = _march(sphere_sdf(c, 0.15), 0.3, 1.5)
    d2 = _march(bulged_sphere_fn(c, 0.15, 0.01), 0.3, 1.5, lipschitz=1.5)
    return d1, d2


def torus_pair(twist=0.02):
    """Torus R = 0.2, r = 0.06 at (0, 0, 0.8) tilted 30 degrees; frame 2 twisted by <= 2 cm."""
    c = (0.0, 0.0, 0.8)
    d1 = _march(torus_fn(c, 0.2, 0.06), 0.3, 1.5)
    d2 = _march(torus_fn(c, 0.2, 0.06, twist=twist), 0.3, 1.5, lipschitz=1.5)
    return d1, d2


def ellipsoid_pair(seed: int, index: int, max_axis: float = 0.2):
    """Random ellipsoid (semi-axes U(0.08, 0.2) m, clipped to max_axis) at (0, 0, 0.8); frame 2
    deformed by three low-frequency sinusoids of total amplitude <= 1.5 cm."""
    u = uniform(seed, 32, offset=32 * index)
    axes = np.minimum(0.08 + 0.12 * u[0:3], max_axis)
    terms = []
    amps = u[3:6]
    amps = 0.015 * amps / max(amps.sum(), 1e-9)
    for t in range(3):
        k = (u[6 + 3 * t: 9 + 3 * t] - 0.5) * 2 * 8.0  # |k| up to ~14 rad/m
        dirv = u[15 + 3 * t: 18 + 3 * t] - 0.5
        dirv = dirv / max(np.linalg.norm(dirv), 1e-9)
        terms.append((float(amps[t]), k, float(2 * math.pi * u[24 + t]), dirv))
    c = (0.0, 0.0, 0.8)
    bound = (c, float(axes.max()) + 0.03)  # the deformation moves the surface by well under 3 cm
    d1 = _march(ellipsoid_fn(c, axes), 0.3, 1.5, lipschitz=2.5, bound=bound)
    d2 = _march(ellipsoid_fn(c, axes, terms), 0.3, 1.5, lipschitz=3.0, bound=bound)
    return d1, d2, axes


# One voxel size for the whole batch (one set of optimiser params).  At 10 mm the smallest
# grid, 48^3, spans 0.48 m, so the largest ellipsoid (semi-axes up to 0.2 m) fits whole.
BATCH_VOXEL = 0.010
BATCH_AXIS_FRAC = 0.45


def batch_edges(count: int = 16, seed: int = 0x5EED, lo: int = 48, hi: int = 160):
    """Cube edges drawn uniformly from lo..hi."""
    u = uniform(seed ^ 0xB47C4, count)
    return [int(lo + min(int(x * (hi - lo + 1)), hi - lo)) for x in u]


def batch_scene(count: int = 16, seed: int = 0x5EED, edges=None, voxel: float = BATCH_VOXEL):
    """Independent volumes of uneven size: volume b is an n_b^3 grid of `voxel` centred on
    (0, 0, 0.8), holding ellipsoid_pair(seed, b).  Semi-axes are clipped to 0.45 of the grid
    extent so the object fits; at BATCH_VOXEL that never binds for edges of 45 or more, so the
    batched workload's ellipsoids keep their full U(0.08, 0.2) m range.
    Returns [(n, min_corner, depth_canonical, depth_live)]."""
    edges = batch_edges(count, seed) if edges is None else list(edges)
    out = []
    for b, n in enumerate(edges):
        ext = n * voxel
        d1, d2, _ = ellipsoid_pair(seed, b, max_axis=BATCH_AXIS_FRAC * ext)
        out.append((n, (-ext / 2, -ext / 2, 0.8 - ext / 2), d1, d2))
    return out


def hinge_frame(t: float, seed: int = 0x5EED, frame: int = 0, noise: bool = True):
    """Two capsules (r = 5 cm, length 0.25 m) joined at a hinge at (0, 0, 0.8); the second
    one swings by 45*t degrees about the camera y axis.  Depth noise N(0, (1 mm * z^2)^2)."""
    r, L = 0.05, 0.25
    h = np.array([0.0, 0.0, 0.8])
    a1 = h + np.array([-L, 0.0, 0.0])
    ang = math.radians(45.0 * t)
    dirv = np.array([math.cos(ang), 0.0, -math.sin(ang)])
    b2 = h + L * dirv
    d = _march(union(capsule_sdf(a1, h, r), capsule_sdf(h, b2, r)), 0.3, 1.5)
    if noise:
        nz = normal(seed, d.size, offset=frame * d.size).reshape(d.shape)
        z = d.astype(np.float64)
        d = np.where(d > 0, z + nz * 0.001 * z * z, 0.0).astype(np.float32)
    return d


def _ray_capsule(dn, a, b, r):
    """Distance along unit rays dn (from the camera centre) to capsule (segment a-b, radius
    r), inf where missed: the cylinder body by the quadratic in the ray parameter, then the
    end-cap spheres where the body hit falls outside the segment."""
    ba = b - a
    oa = -a
    baba = float(ba @ ba)
    bard = dn @ ba
    baoa = float(ba @ oa)
    rdoa = dn @ oa
    oaoa = float(oa @ oa)
    qa = baba - bard * bard
    qb = baba * rdoa - baoa * bard
    qc = baba * oaoa - baoa * baoa - r * r * baba
    h = qb * qb - qa * qc
    t = np.full(dn.shape[0], np.inf)
    ok = (h >= 0.0) & (qa > 1e-12)
    tb = np.where(ok, (-qb - np.sqrt(np.maximum(h, 0.0))) / np.where(qa > 1e-12, qa, 1.0), np.inf)
    y = baoa + tb * bard
    body = ok & (y > 0.0) & (y < baba) & (tb > 0.0)
    t[body] = tb[body]
    # caps: sphere at a (y <= 0) or b (y >= baba)
    for centre in (a, b):
        oc = -centre
        b2 = dn @ oc
        c2 = float(oc @ oc) - r * r
        h2 = b2 * b2 - c2
        ts = np.where(h2 >= 0.0, -b2 - np.sqrt(np.maximum(h2, 0.0)), np.inf)
        ts = np.where(ts > 0.0, ts, np.inf)
        t = np.where(~body, np.minimum(t, ts), t)
    return t


def hinge_frame_exact(t: float, seed: int = 0x5EED, frame: int = 0, noise: bool = True):
    """hinge_frame rendered in closed form (ray-capsule intersection instead of sphere
    tracing): the same scene and noise, a few milliseconds per frame."""
    r, L = 0.05, 0.25
    h = np.array([0.0, 0.0, 0.8])
    a1 = h + np.array([-L, 0.0, 0.0])
    ang = math.radians(45.0 * t)
    b2 = h + L * np.array([math.cos(ang), 0.0, -math.sin(ang)])
    dx, dy = _rays()
    d = np.stack([dx, dy, np.ones_like(dx)], axis=1)
    norm = np.linalg.norm(d, axis=1)
    dn = d / norm[:, None]
    dist = np.minimum(_ray_capsule(dn, a1, h, r), _ray_capsule(dn, h, b2, r))
    z = dist / norm  # camera z of the hit (the ray has unit z per unit of d)
    hit = np.isfinite(z) & (z > 0.3) & (z < 1.5)
    depth = np.where(hit, z, 0.0).reshape(HEIGHT, WIDTH)
    if noise:
        nz = normal(seed, depth.size, offset=frame * depth.size).reshape(depth.shape)
        depth = np.where(depth > 0, depth + nz * 0.001 * depth * depth, 0.0)
    return depth.astype(np.float32)


SEQ_GRID, SEQ_VOXEL = 192, 0.0035
SEQ_MIN = (-0.336, -0.336, 0.414)  # 192 x 3.5 mm = 0.672 m cube around the hinge at z = 0.8


def hinge_sequence(frames: int = 100, seed: int = 0x5EED):
    """The full-sequence workload: `frames` noisy depth frames of the capsule hinge sweeping
    0 to 45 degrees."""
    return [hinge_frame_exact(i / max(frames - 1, 1), seed, i) for i in range(frames)]


def seeded_field(dims, amplitude=0.5, seed=0x5EED):
    """SoA float32 [3, N] displacements uniform in [-amplitude, amplitude) voxels."""
    n = int(dims[0]) * int(dims[1]) * int(dims[2])
    u = uniform(seed, 3 * n)
    return ((u * 2.0 - 1.0) * amplitude).astype(np.float32).reshape(3, n)
