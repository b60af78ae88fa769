"""Pins the FP64 oracle (oracle/kf_oracle.c) before it is trusted as the checker.

Each test is a known-answer test the reference itself states, a closed form derived from
the reference formulas, or an invariant a plausible transcription slip would break.
CPU only.
"""
import math

import dataclasses

import numpy as np
import pytest

import oracle as O


def linear_grid(n=17, voxel=0.5):
    """The reference self-tests' grid: voxel 0.5, bounds [0, 8]^3 -> 17^3, value x + y - z
    (SDF.cpp:526-547)."""
    dims = O.dims_from_bounds((0, 0, 0), (8, 8, 8), voxel)
    assert dims == (n, n, n)
    g = O.OGrid(dims=dims, voxel=voxel, oob=0.048 + 1e-5)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return g, (x + y - z).astype(np.float64).ravel()


def test_get_distance_linear_kat():
    # SDF::testGetDistance, SDF.cpp:550-566: exact to 1e-5 on the interior.
    g, phi = linear_grid()
    for z in range(2, 15):
        for y in range(2, 15, 3):
            for x in range(2, 15, 2):
                for i in range(10):
                    d = i / 10.0
                    f1 = x + d + y - d - z + d
                    f2 = O.get_distance(phi, g, (x + 0.5 + d, y + 0.5 - d, z + 0.5 - d))
                    assert abs(f1 - f2) < 1e-5


def test_get_weight_linear_kat():
    # SDF::testGetWeight, SDF.cpp:606-648 (integer weights x + y - z).
    g, phi = linear_grid()
    w = phi.astype(np.int64)
    for z in range(2, 15, 2):
        for y in range(2, 15, 3):
            for x in range(2, 15, 2):
                for i in range(10):
                    d = float(np.float32(i / 10.0))
                    f1 = x + d + y - d - z + d
                    assert abs(f1 - O.get_weight(w, g, (x + 0.5 + d, y + 0.5 - d, z + 0.5 - d))) < 1e-5


def test_distance_gradient_plain_kat():
    # SDF::testComputeDistanceGradient, SDF.cpp:705-719: gradient (1, 1, -1).
    g, phi = linear_grid()
    for z in range(2, 15, 2):
        for y in range(2, 15, 2):
            for x in range(2, 15, 2):
                for i in range(10):
                    d = float(np.float32(i / 10.0))
                    gr = O.distance_gradient_plain(phi, g, (x + 0.5 + d, y + 0.5 - d, z + 0.5 - d))
                    assert np.allclose(gr, [1, 1, -1], atol=1e-5)


def test_distance_hessian_plain_kat():
    # SDF::testComputeDistanceHessian, SDF.cpp:821-836: linear field -> isZero(1e-4).
    g, phi = linear_grid()
    for z in range(2, 15, 3):
        for y in range(2, 15, 3):
            for x in range(2, 15, 3):
                for i in range(10):
                    d = float(np.float32(i * 0.5 / 10.0))
                    H = O.distance_hessian_plain(phi, g, (x + 0.5 + d, y + 0.5 - d, z + 0.5 - d))
                    assert np.abs(H).max() <= 1e-4


def test_hessian_detects_quadratic():
    # Not a reference KAT: a quadratic field must give its constant Hessian exactly where
    # all 19 samples stay inside one interpolation cell pattern (catches a dropped or
    # swapped sample in the 19-point combination).
    n = 12
    g = O.OGrid(dims=(n, n, n), voxel=1.0)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # Trilinear interpolation of a bilinear xy term is exact, so f = 3xy + 2yz - xz is
    # reproduced exactly and the mixed second derivatives are (3, 2, -1).
    phi = (3.0 * x * y + 2.0 * y * z - 1.0 * x * z).ravel().astype(np.float64)
    H = O.distance_hessian_plain(phi, g, (5.5 + 0.3, 5.5 + 0.3, 5.5 + 0.3), h=0.05)
    assert H[0, 1] == pytest.approx(3.0, abs=1e-9)
    assert H[1, 2] == pytest.approx(2.0, abs=1e-9)
    assert H[0, 2] == pytest.approx(-1.0, abs=1e-9)
    assert np.allclose(np.diag(H), 0.0, atol=1e-9)


def test_jacobian_kat():
    # DisplacementField::testJacobian, DisplacementField.cpp:142-169: F = (x+y+z, x+z, y+z)
    # set for x, y, z < 5 on an 8^3 grid; numerical Jacobian at (2, 2, 2).
    n = 8
    g = O.OGrid(dims=(n, n, n), voxel=0.5)
    u = np.zeros((3, n ** 3))
    for x in range(5):
        for y in range(5):
            for z in range(5):
                i = z * n * n + y * n + x
                u[:, i] = (x + y + z, x + z, y + z)
    J = O.jacobian(u, g, 2, 2, 2)
    assert np.allclose(J, [[1, 1, 1], [1, 0, 1], [0, 1, 1]], atol=1e-12)


def test_killing_energy_linear_closed_form():
    # E_k = ||J||_F^2 + gamma * tr(J J) for a linear field (DisplacementField.cpp:171-193).
    n = 8
    g = O.OGrid(dims=(n, n, n), voxel=0.5)
    A = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [0.2, 0.1, -0.1]])
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    P = np.stack([x.ravel(), y.ravel(), z.ravel()]).astype(np.float64)
    u = A @ P
    gamma = 0.1
    e = O.killing_energy(u, g, 3.0, 3.0, 3.0, gamma=gamma)
    assert e == pytest.approx(np.sum(A * A) + gamma * np.trace(A @ A), rel=1e-12)


def _killing_field(n, fn):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([f(x, y, z).ravel().astype(np.float64) for f in fn])


def test_killing_gradient_closed_forms():
    # Derived by hand from DisplacementField.cpp:305-346 with the trilinear samples of
    # utils.h: pure terms are 2h (u+ + u- - 2u0), mixed terms h^2 (u++ + u-- - u+- - u-+),
    # then the 0.1 clamp and the 1/(4h^2) division.
    n = 8
    g = O.OGrid(dims=(n, n, n), voxel=0.5)
    zero = lambda x, y, z: 0 * x
    # u = (x^2, 0, 0): raw = -2 * 2h * 2 = -0.4 -> clamped to -0.1 -> -10.
    u = _killing_field(n, [lambda x, y, z: x * x, zero, zero])
    for p in [(1, 1, 1), (3, 4, 2), (6, 6, 6)]:
        assert np.allclose(O.killing_gradient2(u, g, *p), [-10, 0, 0], atol=1e-9)
    # u = (a y^2, 0, 0), a < 0.25: unclamped -> -40 a.
    for a in (0.05, 0.2):
        u = _killing_field(n, [lambda x, y, z, a=a: a * y * y, zero, zero])
        assert np.allclose(O.killing_gradient2(u, g, 3, 3, 3), [-40 * a, 0, 0], atol=1e-9)
    # linear u -> 0
    u = _killing_field(n, [lambda x, y, z: 0.1 * x + 0.2 * y, lambda x, y, z: -0.3 * z, zero])
    assert np.allclose(O.killing_gradient2(u, g, 3, 3, 3, gamma=0.1), 0.0, atol=1e-12)
    # gamma couples components through grad div: u = (0, c x y, 0) has a mixed difference
    # u^y_xy = 4 c h^2 (trilinear is exact on a bilinear term), so raw = (-8 gamma c h^2, 0, 0).
    c = 0.4
    u = _killing_field(n, [zero, lambda x, y, z: c * x * y, zero])
    kg = O.killing_gradient2(u, g, 3, 3, 3, gamma=0.1)
    assert kg[0] == pytest.approx(-2 * 0.1 * c, abs=1e-9)  # -2 gamma (4 c h^2) / (4 h^2)
    assert kg[1] == pytest.approx(0.0, abs=1e-12) and kg[2] == pytest.approx(0.0, abs=1e-12)


def test_data_gradient_invariants():
    n = 10
    g = O.OGrid(dims=(n, n, n), voxel=0.008, oob=0.024 + 1e-5)
    p = O.OParams(voxel_size=0.008, max_surface=0.024, clip=0.024, use_levelset=False, use_killing=False)
    rng = np.random.default_rng(1)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a = np.array([0.002, -0.001, 0.003])
    live = (a[0] * x + a[1] * y + a[2] * z - 0.01).ravel().astype(np.float64)
    u = np.zeros((3, n ** 3))
    # Identical SDFs, u = 0 -> zero data gradient (P6).
    for q in [(3, 4, 5), (6, 2, 3)]:
        assert np.allclose(O.energy_gradient(live, live, u, g, p, *q), 0.0, atol=1e-15)
    # Constant shift c -> (c / VoxelSize) * grad (linear field: grad = a exactly).
    c = 0.004
    can = live - c
    for q in [(3, 4, 5), (6, 2, 3)]:
        assert np.allclose(O.energy_gradient(live, can, u, g, p, *q), c / 0.008 * a, rtol=1e-9, atol=1e-15)
    # A constant displacement d moves the data sample to idx + d (SDF.cpp:570-575).
    d = np.array([0.3, -0.2, 0.1])
    ud = np.repeat(d[:, None], n ** 3, axis=1)
    q = (4, 4, 4)
    expect = ((a @ (np.array(q) + d)) - 0.01 - can[4 * n * n + 4 * n + 4]) / 0.008 * a
    assert np.allclose(O.energy_gradient(live, can, ud, g, p, *q), expect, rtol=1e-9, atol=1e-15)
    del rng


def test_levelset_gradient_zero_on_linear():
    # P7: linear phi -> Hessian 0 -> level-set term 0.
    n = 10
    g = O.OGrid(dims=(n, n, n), voxel=0.008, oob=0.024 + 1e-5)
    p = O.OParams(voxel_size=0.008, max_surface=0.024, clip=0.024, use_data=False, use_killing=False)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    live = (0.002 * x - 0.001 * y + 0.003 * z).ravel().astype(np.float64)
    u = np.full((3, n ** 3), 0.2)
    assert np.allclose(O.energy_gradient(live, live, u, g, p, 4, 5, 3), 0.0, atol=1e-12)


def test_displacement_read_wraps_linear_index():
    # DisplacementField.cpp:30-38: x = nx reads voxel (0, y+1, z).
    n = 4
    g = O.OGrid(dims=(n, n, n), voxel=1.0)
    u = np.arange(3 * n ** 3, dtype=np.float64).reshape(3, -1)
    # trilinear at (n - 1 + 0.5, 1, 1): corners (n-1, 1, 1) and (n, 1, 1) -> (0, 2, 1)
    v = O.disp_at_f(u, g, (n - 1 + 0.5, 1.0, 1.0))
    i0 = 1 * n * n + 1 * n + (n - 1)
    i1 = 1 * n * n + 2 * n + 0
    assert np.allclose(v, 0.5 * u[:, i0] + 0.5 * u[:, i1])
    # negative location in (-1, 0): truncation -> base 0, weight negative (extrapolation)
    v = O.disp_at_f(u, g, (-0.25, 0.0, 0.0))
    assert np.allclose(v, 1.25 * u[:, 0] - 0.25 * u[:, 1])


def test_integrate_fronto_parallel_plane():
    # P8: a plane at depth D fills phi = D - Z_c, w = 1 where the pixel lands and the
    # voxel is not more than `clip` behind the surface.
    from killingfusion_amd import scenes
    n = 24
    voxel = 0.01
    mn = (-0.12, -0.12, 0.6)
    g = O.OGrid(dims=(n, n, n), min_corner=mn, voxel=voxel, oob=0.03 + 1e-5)
    D = 0.72
    depth = np.full((scenes.HEIGHT, scenes.WIDTH), D, dtype=np.float32)
    phi, w = O.sdf_init(g)
    O.integrate(phi, w, g, 0.03, depth, scenes.K, scenes.C2W, 0.1, 3.0)
    Dz = float(np.float32(D))
    for z in range(n):
        Zc = voxel * (z + 0.5) + mn[2]
        sl = slice(z * n * n, (z + 1) * n * n)
        if Dz - Zc < -0.03:
            assert (w[sl] == 0).all() and (phi[sl] == g.oob).all()
        else:
            assert (w[sl] == 1).all()
            assert np.allclose(phi[sl], Dz - Zc, atol=1e-15)


def test_integrate_pixel_pick_is_shifted():
    # SDF.cpp:167: col = roundf(u + 0.5), i.e. floor(u) + 1 for u >= 0 -- a one-pixel shift.
    from killingfusion_amd import scenes
    n = 8
    voxel = 0.01
    g = O.OGrid(dims=(n, n, n), min_corner=(0.0, 0.0, 0.5), voxel=voxel, oob=1.0)
    depth = np.zeros((scenes.HEIGHT, scenes.WIDTH), dtype=np.float32)
    # voxel (0, 0, 0) centre projects to u = 525 * 0.005 / 0.505 + 319.5 ~ 324.70
    X, Z = 0.005, 0.505
    ucoord = 525 * X / Z + 319.5
    col = math.floor(ucoord) + 1
    row = math.floor(525 * X / Z + 239.5) + 1
    depth[row, col] = 0.51
    phi, w = O.sdf_init(g)
    O.integrate(phi, w, g, 0.03, depth, scenes.K, scenes.C2W, 0.1, 3.0)
    assert w[0] == 1 and phi[0] == pytest.approx(float(np.float32(0.51)) - Z, abs=1e-12)


def test_fuse_and_warp_invariants():
    n = 9
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    rng = np.random.default_rng(3)
    phi = rng.uniform(-0.03, 0.03, n ** 3)
    w = rng.integers(0, 3, n ** 3).astype(np.int64)
    # P9: fuse(self) keeps phi, doubles w (processTest(1), KillingFusion.cpp:182-188).
    p1, w1 = phi.copy(), w.copy()
    O.fuse(p1, w1, phi.copy(), w.copy())
    assert np.allclose(p1, phi, atol=1e-15) and (w1 == 2 * w).all()
    # empty target -> copy where the source has weight
    p2, w2 = O.sdf_init(g)
    O.fuse(p2, w2, phi, w)
    m = w > 0
    assert (p2[m] == phi[m]).all() and (w2 == w).all() and (p2[~m] == g.oob).all()
    # P10: warp with u = 0 keeps nodes.
    p3, w3 = phi.copy(), w.copy()
    O.warp(p3, w3, np.zeros((3, n ** 3)), g)
    assert np.allclose(p3, phi, atol=1e-15) and (w3 == w).all()
    # integer shift by +1 in x reads the next node; the last column reads out of grid.
    p4, w4 = phi.copy(), w.copy()
    u = np.zeros((3, n ** 3))
    u[0] = 1.0
    O.warp(p4, w4, u, g)
    P = phi.reshape(n, n, n)
    Q = p4.reshape(n, n, n)
    assert np.allclose(Q[:, :, :-1], P[:, :, 1:], atol=1e-15)
    assert np.allclose(Q[:, :, -1], g.oob)


def test_fuse_warped_zero_field_equals_fuse():
    n = 9
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    rng = np.random.default_rng(4)
    phi = rng.uniform(-0.03, 0.03, n ** 3)
    w = rng.integers(0, 3, n ** 3).astype(np.int64)
    a, aw = O.sdf_init(g)
    a[:] = rng.uniform(-0.03, 0.03, n ** 3)
    aw[:] = rng.integers(0, 2, n ** 3)
    b, bw = a.copy(), aw.copy()
    O.fuse(a, aw, phi, w)
    O.fuse_warped(b, bw, phi, w, np.zeros((3, n ** 3)), g)
    assert np.allclose(a, b, atol=1e-15) and (aw == bw).all()


def test_mc_table_structure():
    # P11: for all 256 configurations the case's triangles use exactly the cut edges,
    # and the edge mask is the set of edges with differently signed corners.
    corners = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for ci in range(256):
        em = O.mc_edge_mask(ci)
        expect = sum(1 << e for e, (a, b) in enumerate(corners) if ((ci >> a) & 1) != ((ci >> b) & 1))
        assert em == expect
        row = O.mc_case(ci)
        case = [e for e in row if e >= 0]
        assert all(e == -1 for e in row[len(case):])
        assert len(case) % 3 == 0 and len(case) // 3 <= 5
        used = set(case)
        assert used == {e for e in range(12) if em >> e & 1}


def test_mc_sphere_is_closed_and_deterministic():
    # A sphere well inside the grid must give a watertight soup: every edge (as a pair of
    # vertex positions) belongs to exactly two triangles.  A wrong table entry opens holes.
    n = 20
    g = O.OGrid(dims=(n, n, n), voxel=1.0, oob=4.0 + 1e-5)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    phi = (np.sqrt((x - 9.3) ** 2 + (y - 10.1) ** 2 + (z - 9.7) ** 2) - 5.2).ravel()
    tris = O.marching_cubes(phi, g, 100.0, 100.0)
    assert len(tris) > 100
    key = lambda v: tuple(np.round(v, 7))
    from collections import Counter
    cnt = Counter()
    for t in tris:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = tuple(sorted((key(t[a]), key(t[b]))))
            cnt[e] += 1
    assert set(cnt.values()) == {2}
    again = O.marching_cubes(phi, g, 100.0, 100.0)
    assert np.array_equal(tris, again)


def test_mc_skip_rule_and_y_flip():
    # SDF.cpp:405-408: a cell is skipped when its (x, y, z) corner is > Max or < -clip.
    n = 6
    g = O.OGrid(dims=(n, n, n), voxel=1.0, oob=3.0)
    phi = np.full(n ** 3, 1.0)
    phi[2 * n * n + 2 * n + 2] = -1.0  # one inside voxel
    t_all = O.marching_cubes(phi, g, 2.0, 2.0)
    t_skip = O.marching_cubes(phi, g, 0.5, 2.0)  # corner values 1.0 > 0.5: those cells skip
    assert len(t_all) == 8 and len(t_skip) == 1  # only the cell whose own corner is -1
    # vertices: positions v/(n/2) - 1 with y negated
    v = t_skip.reshape(-1, 3)
    assert np.all((v[:, 1] >= -1e-12) & (v[:, 1] <= 1 / 3 + 1e-12))  # y in [-1/3, 0], negated


def test_deform_jacobi_matches_naive_restatement():
    # P12: one Jacobi iteration of the C oracle equals a pure-Python re-derivation of
    # KillingFusion.cpp:251-351 on a tiny grid (independent transcription).
    n = 6
    rng = np.random.default_rng(7)
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, omega_k=0.5, gamma=0.1,
                  omega_s=0.2, max_iters=1)
    live = rng.uniform(-0.03, 0.03, n ** 3)
    can = rng.uniform(-0.03, 0.03, n ** 3)
    u = rng.uniform(-0.4, 0.4, (3, n ** 3))
    un, st = O.deform_jacobi(live, can, u, g, p)
    ref = _naive_jacobi(live, can, u, n, p)
    assert np.allclose(un, ref, atol=1e-12, rtol=1e-10)


def test_trust_region_matches_naive_restatement():
    # The trust-region step of the all-voxels mode (KillingFusion.cpp:288-314, UseTrustStrategy
    # with the data energy alone): one iteration of the C oracle equals an independent Python
    # transcription, including the reference's `break`, which leaves the rest of an x-row
    # without a step once a voxel's step size falls below 1e-7.  The random fields make some
    # voxels shrink their step and some rows break.
    n = 6
    rng = np.random.default_rng(21)
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, use_levelset=False, use_killing=False,
                  use_trust=True, max_iters=1)
    live = rng.uniform(-0.03, 0.03, n ** 3)
    can = rng.uniform(-0.03, 0.03, n ** 3)
    u = rng.uniform(-0.4, 0.4, (3, n ** 3))
    un, st = O.deform_jacobi(live, can, u, g, p)
    ref, shrunk, breaks, skipped = _naive_trust(live, can, u, n, p)
    assert np.array_equal(un, ref)
    assert shrunk > 0 and breaks > 0 and skipped > 0, (shrunk, breaks, skipped)
    # without the trust region (or with another energy on) the step is the plain one
    plain, _ = O.deform_jacobi(live, can, u, g, dataclasses.replace(p, use_trust=False))
    assert not np.array_equal(plain, ref)
    other, _ = O.deform_jacobi(live, can, u, g, dataclasses.replace(p, use_levelset=True))
    plain_ls, _ = O.deform_jacobi(live, can, u, g, dataclasses.replace(p, use_levelset=True, use_trust=False))
    assert np.array_equal(other, plain_ls)


def _naive_trust(live, can, u, n, p):
    """One all-voxels iteration with the data energy alone and the trust region, transcribed
    from KillingFusion.cpp:251-319 and SDF.cpp:503-524 (trilinear read with truncating casts);
    returns the field and the counts of shrunk steps, row breaks and voxels they skipped."""
    N = n ** 3
    oob = p.max_surface + p.epsilon

    def phi_at(arr, x, y, z):
        if 0 <= x < n and 0 <= y < n and 0 <= z < n:
            return arr[z * n * n + y * n + x]
        return oob

    def getdist(arr, l):
        t = [c - 0.5 for c in l]
        b = [int(c) for c in t]
        f = [t[k] - b[k] for k in range(3)]
        v = {(dx, dy, dz): phi_at(arr, b[0] + dx, b[1] + dy, b[2] + dz)
             for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)}
        L = lambda a, c, x: a * (1 - x) + c * x
        s = L(L(v[0, 0, 0], v[1, 0, 0], f[0]), L(v[0, 1, 0], v[1, 1, 0], f[0]), f[1])
        t2 = L(L(v[0, 0, 1], v[1, 0, 1], f[0]), L(v[0, 1, 1], v[1, 1, 1], f[0]), f[1])
        return L(s, t2, f[2])

    def dat(x, y, z):
        i = z * n * n + y * n + x
        return u[:, i] if 0 <= i < N else np.zeros(3)

    def dispf(l):
        b = [int(c) for c in l]
        f = [l[k] - b[k] for k in range(3)]
        L = lambda a, c, x: a * (1 - x) + c * x
        c = {(dx, dy, dz): dat(b[0] + dx, b[1] + dy, b[2] + dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)}
        s = L(L(c[0, 0, 0], c[1, 0, 0], f[0]), L(c[0, 1, 0], c[1, 1, 0], f[0]), f[1])
        t2 = L(L(c[0, 0, 1], c[1, 0, 1], f[0]), L(c[0, 1, 1], c[1, 1, 1], f[0]), f[1])
        return L(s, t2, f[2])

    def distf(l):
        d = dispf(l)
        return getdist(live, [l[k] + d[k] + 0.5 for k in range(3)])

    h = p.h
    out = u.copy()
    shrunk = breaks = skipped = 0
    E = np.eye(3)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                i = z * n * n + y * n + x
                psi = getdist(live, [x + u[0, i] + 0.5, y + u[1, i] + 0.5, z + u[2, i] + 0.5])
                if psi > p.max_surface - p.epsilon or psi < -p.clip:
                    continue
                gl = np.array([x + 0.5, y + 0.5, z + 0.5])
                gr = np.array([distf(gl + h * E[a]) - distf(gl - h * E[a]) for a in range(3)]) / (2 * h)
                dg = (psi - can[i]) / p.voxel_size * gr
                a = p.alpha
                du = -a * dg
                dest, prev = can[i], psi
                while True:
                    src = getdist(live, [x + u[0, i] + du[0] + 0.5, y + u[1, i] + du[1] + 0.5,
                                         z + u[2, i] + du[2] + 0.5])
                    if abs(src - dest) - abs(prev - dest) > 0:
                        a /= 1.5
                        du = -a * dg
                        shrunk += 1
                        if not a > 1e-7:
                            break
                    else:
                        break
                if a < 1e-7:
                    breaks += 1
                    skipped += n - 1 - x
                    break
                out[:, i] = u[:, i] + du
    return out, shrunk, breaks, skipped


def _naive_jacobi(live, can, u, n, p):
    N = n ** 3
    oob = p.max_surface + p.epsilon

    def phi_at(arr, x, y, z):
        if 0 <= x < n and 0 <= y < n and 0 <= z < n:
            return arr[z * n * n + y * n + x]
        return oob

    def getdist(arr, l):
        t = [c - 0.5 for c in l]
        b = [int(c) for c in t]
        f = [t[k] - b[k] for k in range(3)]
        v = {}
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    v[dx, dy, dz] = phi_at(arr, b[0] + dx, b[1] + dy, b[2] + dz)
        L = lambda a, c, x: a * (1 - x) + c * x
        s = L(L(v[0, 0, 0], v[1, 0, 0], f[0]), L(v[0, 1, 0], v[1, 1, 0], f[0]), f[1])
        t2 = L(L(v[0, 0, 1], v[1, 0, 1], f[0]), L(v[0, 1, 1], v[1, 1, 1], f[0]), f[1])
        return L(s, t2, f[2])

    def dat(x, y, z):
        i = z * n * n + y * n + x
        if i < 0 or i >= N:
            return np.zeros(3)
        return u[:, i]

    def dispf(l):
        b = [int(c) for c in l]
        f = [l[k] - b[k] for k in range(3)]
        L = lambda a, c, x: a * (1 - x) + c * x
        c = {(dx, dy, dz): dat(b[0] + dx, b[1] + dy, b[2] + dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)}
        s = L(L(c[0, 0, 0], c[1, 0, 0], f[0]), L(c[0, 1, 0], c[1, 1, 0], f[0]), f[1])
        t2 = L(L(c[0, 0, 1], c[1, 0, 1], f[0]), L(c[0, 1, 1], c[1, 1, 1], f[0]), f[1])
        return L(s, t2, f[2])

    def distf(l):
        d = dispf(l)
        return getdist(live, [l[k] + d[k] + 0.5 for k in range(3)])

    h = p.h
    out = u.copy()
    for z in range(n):
        for y in range(n):
            for x in range(n):
                i = z * n * n + y * n + x
                psi = getdist(live, [x + u[0, i] + 0.5, y + u[1, i] + 0.5, z + u[2, i] + 0.5])
                if psi > p.max_surface - p.epsilon or psi < -p.clip:
                    continue
                gl = np.array([x + 0.5, y + 0.5, z + 0.5])
                E = np.eye(3)
                gr = np.array([distf(gl + h * E[a]) - distf(gl - h * E[a]) for a in range(3)]) / (2 * h)
                dg = (psi - can[i]) / p.voxel_size * gr
                f = lambda o: distf(gl + h * np.array(o, dtype=float))
                f0 = f((0, 0, 0))
                H = np.zeros((3, 3))
                for a in range(3):
                    e2 = np.zeros(3)
                    e2[a] = 2
                    H[a, a] = f(e2) - 2 * f0 + f(-e2)
                for a, b in ((0, 1), (0, 2), (1, 2)):
                    ea, eb = E[a], E[b]
                    H[a, b] = H[b, a] = f(ea + eb) + f(-ea - eb) - f(ea - eb) - f(-ea + eb)
                H /= 4 * h * h
                nn = np.linalg.norm(gr)
                lg = H @ gr * (nn - 1) / (nn + p.epsilon) * p.omega_s
                k = lambda o: dispf(np.array([x, y, z], dtype=float) + h * np.array(o, dtype=float))
                k0 = k((0, 0, 0))
                D2 = {}
                for a in range(3):
                    e2 = np.zeros(3)
                    e2[a] = 2
                    D2[a, a] = k(e2) - 2 * k0 + k(-e2)
                for a, b in ((0, 1), (0, 2), (1, 2)):
                    ea, eb = E[a], E[b]
                    D2[a, b] = D2[b, a] = k(ea + eb) + k(-ea - eb) - k(ea - eb) - k(-ea + eb)
                lap = D2[0, 0] + D2[1, 1] + D2[2, 2]
                div = np.array([sum(D2[a, b][b] for b in range(3)) for a in range(3)])
                kg = -2 * lap - 2 * p.gamma * div
                if np.linalg.norm(kg) > p.killing_clamp:
                    kg = kg / np.linalg.norm(kg) * p.killing_clamp
                kg = kg / (4 * h * h) * p.omega_k
                out[:, i] = u[:, i] - p.alpha * (dg + kg + lg)
    return out


def test_adjoint_gradient_matches_fd_of_energy():
    # P13: the oracle's analytic adjoint equals central differences of its energy (FP64).
    n = 10
    rng = np.random.default_rng(11)
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, omega_k=0.5, gamma=0.1, omega_s=0.2)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    live = (0.01 * (np.sqrt((x - 4.6) ** 2 + (y - 5.2) ** 2 + (z - 4.9) ** 2) - 3.0)).ravel()
    can = live + rng.uniform(-0.002, 0.002, n ** 3)
    u = rng.uniform(-0.45, 0.45, (3, n ** 3))
    grad, mask = O.adj_gradient(live, can, u, g, p)
    eps = 1e-6
    checked = 0
    for i in rng.choice(n ** 3, 150, replace=False):
        for a in range(3):
            up, um = u.copy(), u.copy()
            up[a, i] += eps
            um[a, i] -= eps
            _, _, ap = O.adj_energy(live, can, up, g, p, per_voxel=True)
            _, _, am = O.adj_energy(live, can, um, g, p, per_voxel=True)
            fd = (O.adj_energy(live, can, up, g, p) - O.adj_energy(live, can, um, g, p)) / (2 * eps)
            # skip where the perturbation moves the sample across a cell face or flips the mask
            q = np.array([x.ravel()[i], y.ravel()[i], z.ravel()[i]])[a] + u[a, i]
            if abs(q - round(q)) < 10 * eps or ap != am:
                continue
            assert fd == pytest.approx(grad[a, i], rel=1e-5, abs=1e-6)
            checked += 1
    assert checked > 300


def test_adjoint_killing_equals_reference_without_clamp():
    # With gamma = 0, no clamp hit and samples inside their cells, the reference Killing
    # gradient (DisplacementField.cpp:301-347) is (1/(2h)) * (-2 Lap u) and the adjoint's
    # Killing part is -2 Lap u: their ratio is exactly 1/(2h) = 10 in the interior.
    n = 10
    rng = np.random.default_rng(5)
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=10.0)
    p = O.OParams(voxel_size=0.01, max_surface=100.0, clip=100.0, use_data=False, use_levelset=False,
                  omega_k=1.0, gamma=0.0, killing_clamp=1e9)
    u = rng.uniform(-0.01, 0.01, (3, n ** 3))
    live = np.zeros(n ** 3)
    grad, _ = O.adj_gradient(live, live, u, g, p)
    for (x, y, z) in [(3, 4, 5), (5, 5, 5), (2, 7, 3)]:
        i = z * n * n + y * n + x
        kref = O.killing_gradient2(u, g, x, y, z, h=0.05, gamma=0.0, clamp=1e9)
        assert np.allclose(kref, grad[:, i] / (2 * 0.05), rtol=1e-9, atol=1e-12)


def test_adjoint_deform_steps_equal_dense_gradient_steps():
    # kfo_adj_deform evaluates the gradient only where a step uses it (active voxels; a voxel
    # with u_j = 0 and an out-of-band live[j] is inactive without sampling).  Its iterations
    # must equal Jacobi steps taken with the dense kfo_adj_gradient: u <- u - alpha mask grad.
    n = 12
    rng = np.random.default_rng(5)
    g = O.OGrid(dims=(n, n + 1, n - 1), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, omega_k=0.5, gamma=0.1, omega_s=0.2,
                  max_iters=3)
    z, y, x = np.meshgrid(np.arange(n - 1), np.arange(n + 1), np.arange(n), indexing="ij")
    live = (0.01 * (np.sqrt((x - 5.6) ** 2 + (y - 6.2) ** 2 + (z - 4.9) ** 2) - 3.5)).ravel()
    live = np.where(np.abs(live) < 0.03, live, 0.03 + 1e-5)  # truncated band, untouched outside
    can = live + rng.uniform(-0.002, 0.002, live.size)
    for u0 in (np.zeros((3, g.n)), np.where(rng.random((3, g.n)) < 0.3, rng.uniform(-0.4, 0.4, (3, g.n)), 0.0)):
        u = u0.copy()
        for _ in range(3):
            grad, mask = O.adj_gradient(live, can, u, g, p)
            u = u - p.alpha * grad * mask
        ud, st = O.adj_deform(live, can, u0, g, p)
        assert st["iters"] == 3 and st["active"] == int(mask.sum()) > 50
        assert np.array_equal(ud, u)


def test_voxels_outside_the_start_list_never_move():
    # The active-set sweep lists C = {j : u_j has a nonzero bit pattern, or live[j] in the band}
    # once per deform call.  Dense Jacobi steps (every voxel's gradient, applied where the step's
    # band mask is set) must leave every voxel outside C at exactly +0, never mark it active,
    # and the energy of every iterate must be carried by C and the voxels just below C
    # (i with some i + e_b in C) -- the set the active-set energy pass visits.
    n = 20
    rng = np.random.default_rng(9)
    g = O.OGrid(dims=(n, n + 1, n - 1), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, omega_k=0.5, gamma=0.1, omega_s=0.2)
    z, y, x = np.meshgrid(np.arange(n - 1), np.arange(n + 1), np.arange(n), indexing="ij")
    live = (0.01 * (np.sqrt((x - 9.6) ** 2 + (y - 10.2) ** 2 + (z - 9.1) ** 2) - 4.5)).ravel()
    live = np.where(np.abs(live) < 0.03, live, 0.03 + 1e-5)
    can = live + rng.uniform(-0.002, 0.002, live.size)
    u0 = np.zeros((3, g.n))
    box = (z < 3) & (y > 15) & (x > 14)  # away from the surface: moved only by the starting field
    u0[:, box.ravel()] = rng.uniform(-0.3, 0.3, (3, int(box.sum())))
    hi, lo = p.max_surface - p.epsilon, -p.clip
    listed = (u0 != 0).any(axis=0) | ~((live > hi) | (live < lo))
    L3 = listed.reshape(n - 1, n + 1, n)
    below = np.zeros_like(L3)
    below[:, :, :-1] |= L3[:, :, 1:]
    below[:, :-1, :] |= L3[:, 1:, :]
    below[:-1, :, :] |= L3[1:, :, :]
    carried = (L3 | below).ravel()
    assert 0 < listed.sum() < g.n // 3
    u = u0.copy()
    for it in range(12):
        grad, mask = O.adj_gradient(live, can, u, g, p)
        assert not (mask.astype(bool) & ~listed).any(), it
        u = u - p.alpha * grad * mask
        assert not u[:, ~listed].any() and not np.signbit(u[:, ~listed]).any(), it
        _, ev, _ = O.adj_energy(live, can, u, g, p, per_voxel=True)
        assert not ev[~carried].any(), it
        assert ev[carried & ~listed].any()  # the terms just below C are not all zero


def test_warp_leaves_unmoved_voxels_alone():
    # SDF::update (SDF.cpp:300-328) at a voxel with u = 0 reads its own cell with zero weights:
    # phi and weight come back unchanged for finite phi (a -0 phi may come back as +0), which
    # is what lets the warp run over the voxels with u != 0 only.
    n = 10
    rng = np.random.default_rng(3)
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    phi = rng.uniform(-0.04, 0.04, g.n)
    phi[[5, 77, 400]] = -0.0
    w = rng.integers(0, 9, g.n).astype(np.int64)
    u = np.where(rng.random((3, g.n)) < 0.1, rng.uniform(-1.5, 1.5, (3, g.n)), 0.0)
    moved = (u != 0).any(axis=0)
    wp, ww = phi.copy(), w.copy()
    O.warp(wp, ww, u, g)
    keep = ~moved & ~((phi == 0) & np.signbit(phi))
    assert np.array_equal(wp[keep], phi[keep]) and np.array_equal(ww[~moved], w[~moved])
    assert (wp[moved] != phi[moved]).any()


def test_adjoint_local_energy_carries_every_term_of_a_voxel():
    # kfo_adj_energy_local(x, y, z) must change exactly as the total energy does when u at
    # (x, y, z) moves (including the zero-padding ghost terms on low faces).
    n = 9
    rng = np.random.default_rng(21)
    g = O.OGrid(dims=(n, n, n), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, omega_k=0.5, gamma=0.1, omega_s=0.2)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    live = (0.01 * (np.sqrt((x - 4.1) ** 2 + (y - 4.6) ** 2 + (z - 3.9) ** 2) - 2.5)).ravel()
    can = live + rng.uniform(-0.002, 0.002, n ** 3)
    u = rng.uniform(-0.4, 0.4, (3, n ** 3))
    picks = [(0, 0, 0), (0, 4, 4), (8, 8, 8), (4, 0, 7)] + [tuple(rng.integers(0, n, 3)) for _ in range(20)]
    for (px, py, pz) in picks:
        i = (pz * n + py) * n + px
        for a in range(3):
            up, um = u.copy(), u.copy()
            up[a, i] += 0.05
            um[a, i] -= 0.03
            dl = O.adj_energy_local(live, can, up, g, p, px, py, pz) - O.adj_energy_local(live, can, um, g, p, px, py, pz)
            dt = O.adj_energy(live, can, up, g, p) - O.adj_energy(live, can, um, g, p)
            assert dl == pytest.approx(dt, rel=1e-9, abs=1e-9)


def test_gauss_seidel_serial_equals_a_transcription_of_the_per_voxel_loop():
    # KillingFusion.cpp:353-396 run in z -> y -> x order: skip test on the displaced distance
    # (symmetric +-(Max - eps)), then repeated in-place steps at that voxel until the step or
    # the gradient is below the threshold or the iteration cap is hit.  An independent Python
    # transcription over the oracle's own per-voxel pieces (get_distance, energy_gradient).
    import math
    g = O.OGrid(dims=(6, 5, 4), voxel=0.01, oob=0.03 + 1e-5)
    p = O.OParams(voxel_size=0.01, max_surface=0.03, clip=0.03, alpha=0.1, omega_k=0.5, gamma=0.1, omega_s=0.2,
                  max_iters=5, gs_threshold=1e-4)
    rng = np.random.default_rng(17)
    z, y, x = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    live = (0.01 * (x + 0.5 * y - 0.3 * z - 2.2)).ravel().astype(np.float64)
    can = live + rng.uniform(-0.003, 0.003, live.size)
    u0 = rng.uniform(-0.3, 0.3, (3, live.size))
    uo, _ = O.deform_gs_serial(live, can, u0, g, p)
    u = u0.copy()
    nx, ny = 6, 5
    visited = 0
    for zz in range(4):
        for yy in range(5):
            for xx in range(6):
                i = (zz * ny + yy) * nx + xx
                # SDF::getDistance(idx, disp), SDF.cpp:570-575: idx + u + 0.5
                d = O.get_distance(live, g, (xx + u[0, i] + 0.5, yy + u[1, i] + 0.5, zz + u[2, i] + 0.5))
                if d > p.max_surface - p.epsilon or d < -p.max_surface + p.epsilon:
                    continue
                visited += 1
                it = 0
                while True:
                    gr = O.energy_gradient(live, can, u, g, p, xx, yy, zz)
                    du = [-p.alpha * gr[0], -p.alpha * gr[1], -p.alpha * gr[2]]
                    for a in range(3):
                        u[a, i] = u[a, i] + du[a]
                    if math.sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]) <= p.gs_threshold:
                        break
                    it += 1
                    if not (math.sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]) > p.gs_threshold
                            and it < p.max_iters):
                        break
    assert visited > 20
    assert np.array_equal(uo, u)


def test_fuse_nearest_surface_mode():
    # FUSE_BY_MERGE = false (config.cpp:7, SDF.cpp:226-237 and :285-295): a voxel the other SDF
    # never saw (w2 = 0) is left alone, an empty one takes the other's value and weight, and
    # otherwise the value with the smaller magnitude wins with its own weight (ties take the
    # other's: the test is |d1| < |d2|).
    phi1 = np.array([0.5, 0.1, 0.02, -0.03, 0.04, -0.04])
    w1 = np.array([3, 0, 2, 2, 5, 5], dtype=np.int64)
    phi2 = np.array([0.9, -0.2, 0.05, 0.01, -0.04, 0.04])
    w2 = np.array([0, 4, 1, 1, 7, 7], dtype=np.int64)
    a, aw = phi1.copy(), w1.copy()
    O.fuse(a, aw, phi2, w2, merge=False)
    assert a.tolist() == [0.5, -0.2, 0.02, 0.01, -0.04, 0.04]
    assert aw.tolist() == [3, 4, 2, 1, 7, 7]
    # the warped variant with a zero field reads the other SDF at the nodes: the same result
    n = 6
    g = O.OGrid(dims=(n, 1, 1), voxel=1.0, oob=1.0)
    b, bw = phi1.copy(), w1.copy()
    O.fuse_warped(b, bw, phi2, w2, np.zeros((3, n)), g, merge=False)
    assert b.tolist() == a.tolist() and bw.tolist() == aw.tolist()
    # and the merge mode is the weighted average of the same inputs
    c, cw = phi1.copy(), w1.copy()
    O.fuse(c, cw, phi2, w2, merge=True)
    assert c[2] == (2 * 0.02 + 1 * 0.05) / 3 and cw.tolist() == [3, 4, 3, 3, 12, 12]
