"""FP64 oracle of KillingFusion's per-frame hot path (ctypes over oracle/build/libkf_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py`` -- as the checker or as the timed CPU baseline, never
as part of the shipped path.  The C source (oracle/kf_oracle.c) cites the reference
file:line of every function it restates.

Arrays: phi float64 [N] (x fastest), weights int64 [N], displacements float64 SoA [3, N]
in voxel units.  Grid dims are (nx, ny, nz).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libkf_oracle.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class Grid(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("min", C.c_double * 3), ("voxel", C.c_double), ("oob", C.c_double)]


class Params(C.Structure):
    _fields_ = [
        ("voxel_size", C.c_double), ("max_surface", C.c_double), ("clip", C.c_double),
        ("epsilon", C.c_double), ("alpha", C.c_double), ("omega_k", C.c_double),
        ("gamma", C.c_double), ("omega_s", C.c_double), ("h", C.c_double),
        ("killing_clamp", C.c_double), ("use_data", C.c_int32), ("use_levelset", C.c_int32),
        ("use_killing", C.c_int32), ("max_iters", C.c_int32), ("stop_max_update", C.c_double),
        ("gs_threshold", C.c_double), ("use_trust", C.c_int32), ("pad", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("diverged_index", C.c_int64), ("max_update", C.c_double),
                ("energy", C.c_double), ("active", C.c_int64)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        g = C.POINTER(Grid)
        p = C.POINTER(Params)
        s = C.POINTER(Stats)
        sig = {
            "kfo_dims_from_bounds": (None, [_dp, _dp, C.c_double, C.POINTER(C.c_int32)]),
            "kfo_distance_at_index": (C.c_double, [_dp, g, C.c_int, C.c_int, C.c_int]),
            "kfo_get_distance": (C.c_double, [_dp, g, _dp]),
            "kfo_get_weight": (C.c_double, [_i64p, g, _dp]),
            "kfo_disp_at_f": (None, [_dp, g, _dp, _dp]),
            "kfo_distance_idx_disp": (C.c_double, [_dp, _dp, g, C.c_int, C.c_int, C.c_int]),
            "kfo_distance_gradient_disp": (None, [_dp, _dp, g, C.c_int, C.c_int, C.c_int, C.c_double, _dp]),
            "kfo_distance_hessian_disp": (None, [_dp, _dp, g, C.c_int, C.c_int, C.c_int, C.c_double, _dp]),
            "kfo_distance_gradient_plain": (C.c_int, [_dp, g, _dp, _dp]),
            "kfo_distance_hessian_plain": (C.c_int, [_dp, g, _dp, C.c_double, _dp]),
            "kfo_jacobian": (None, [_dp, g, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
            "kfo_killing_energy": (C.c_double, [_dp, g, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
            "kfo_killing_gradient2": (None, [_dp, g, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp]),
            "kfo_data_gradient": (None, [_dp, _dp, _dp, g, p, C.c_int, C.c_int, C.c_int, _dp]),
            "kfo_levelset_gradient": (None, [_dp, _dp, g, p, C.c_int, C.c_int, C.c_int, _dp]),
            "kfo_energy_gradient": (None, [_dp, _dp, _dp, g, p, C.c_int, C.c_int, C.c_int, _dp]),
            "kfo_energy_gradient_grid": (None, [_dp, _dp, _dp, g, p, C.c_int, _dp]),
            "kfo_deform_jacobi": (C.c_int, [_dp, _dp, _dp, g, p, s]),
            "kfo_deform_gs_serial": (C.c_int, [_dp, _dp, _dp, g, p, s]),
            "kfo_integrate": (None, [_dp, _i64p, g, C.c_double, C.POINTER(C.c_float), C.c_int, C.c_int,
                                     _dp, _dp, C.c_double, C.c_double]),
            "kfo_invert4": (C.c_int, [_dp, _dp]),
            "kfo_warp": (None, [_dp, _i64p, _dp, g]),
            "kfo_fuse": (None, [_dp, _i64p, _dp, _i64p, C.c_int64]),
            "kfo_fuse_warped": (None, [_dp, _i64p, _dp, _i64p, _dp, g]),
            "kfo_fuse_mode": (None, [_dp, _i64p, _dp, _i64p, C.c_int64, C.c_int]),
            "kfo_fuse_warped_mode": (None, [_dp, _i64p, _dp, _i64p, _dp, g, C.c_int]),
            "kfo_marching_cubes": (C.c_int64, [_dp, g, C.c_double, C.c_double, C.c_double, _dp, C.c_int64]),
            "kfo_mc_edge_mask": (C.c_int, [C.c_int]),
            "kfo_mc_case": (None, [C.c_int, C.POINTER(C.c_int32)]),
            "kfo_adj_energy": (C.c_double, [_dp, _dp, _dp, g, p, _dp, _i64p]),
            "kfo_adj_gradient": (None, [_dp, _dp, _dp, g, p, _dp, _u8p]),
            "kfo_adj_energy_local": (C.c_double, [_dp, _dp, _dp, g, p, C.c_int, C.c_int, C.c_int]),
            "kfo_adj_deform": (C.c_int, [_dp, _dp, _dp, g, p, s]),
            "kfo_num_threads": (C.c_int, []),
            "kfo_set_num_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


@dataclass
class OGrid:
    """Grid geometry: dims (nx, ny, nz), min corner (m), voxel (m), out-of-grid phi value."""
    dims: tuple
    min_corner: tuple = (0.0, 0.0, 0.0)
    voxel: float = 0.012
    oob: float = 0.048 + 1e-5

    @property
    def n(self) -> int:
        return int(self.dims[0]) * int(self.dims[1]) * int(self.dims[2])

    def c(self) -> Grid:
        g = Grid()
        for a in range(3):
            g.n[a] = int(self.dims[a])
            g.min[a] = float(self.min_corner[a])
        g.voxel = float(self.voxel)
        g.oob = float(self.oob)
        return g


@dataclass
class OParams:
    """Energy/optimiser parameters; defaults are the reference's config.cpp:4-57."""
    voxel_size: float = 0.012
    max_surface: float = 0.048
    clip: float = 0.048
    epsilon: float = 1e-5
    alpha: float = 0.025
    omega_k: float = 0.04
    gamma: float = 0.0
    omega_s: float = 0.1
    h: float = 0.05
    killing_clamp: float = 0.1
    use_data: bool = True
    use_levelset: bool = True
    use_killing: bool = True
    max_iters: int = 1024
    stop_max_update: float = 0.0
    gs_threshold: float = 1e-6
    use_trust: bool = False
    pad: int = 0

    def c(self) -> Params:
        p = Params()
        for f in Params._fields_:
            setattr(p, f[0], getattr(self, f[0]))
        return p


def dims_from_bounds(mn, mx, voxel):
    out = (C.c_int32 * 3)()
    a = np.array(mn, dtype=np.float64)
    b = np.array(mx, dtype=np.float64)
    lib().kfo_dims_from_bounds(_d(a), _d(b), voxel, out)
    return tuple(out)


def get_distance(phi, g: OGrid, loc):
    phi = _c64(phi)
    l = np.array(loc, dtype=np.float64)
    return lib().kfo_get_distance(_d(phi), C.byref(g.c()), _d(l))


def get_weight(w, g: OGrid, loc):
    w = np.ascontiguousarray(w, dtype=np.int64)
    l = np.array(loc, dtype=np.float64)
    return lib().kfo_get_weight(w.ctypes.data_as(_i64p), C.byref(g.c()), _d(l))


def disp_at_f(u, g: OGrid, loc):
    u = _c64(u)
    l = np.array(loc, dtype=np.float64)
    o = np.zeros(3)
    lib().kfo_disp_at_f(_d(u), C.byref(g.c()), _d(l), _d(o))
    return o


def distance_gradient_plain(phi, g: OGrid, loc):
    phi = _c64(phi)
    l = np.array(loc, dtype=np.float64)
    o = np.zeros(3)
    lib().kfo_distance_gradient_plain(_d(phi), C.byref(g.c()), _d(l), _d(o))
    return o


def distance_hessian_plain(phi, g: OGrid, loc, h=0.05):
    phi = _c64(phi)
    l = np.array(loc, dtype=np.float64)
    o = np.zeros(9)
    lib().kfo_distance_hessian_plain(_d(phi), C.byref(g.c()), _d(l), h, _d(o))
    return o.reshape(3, 3)


def distance_gradient_disp(phi, u, g: OGrid, x, y, z, h=0.05):
    phi, u = _c64(phi), _c64(u)
    o = np.zeros(3)
    lib().kfo_distance_gradient_disp(_d(phi), _d(u), C.byref(g.c()), x, y, z, h, _d(o))
    return o


def distance_hessian_disp(phi, u, g: OGrid, x, y, z, h=0.05):
    phi, u = _c64(phi), _c64(u)
    o = np.zeros(9)
    lib().kfo_distance_hessian_disp(_d(phi), _d(u), C.byref(g.c()), x, y, z, h, _d(o))
    return o.reshape(3, 3)


def jacobian(u, g: OGrid, x, y, z, h=0.05):
    u = _c64(u)
    o = np.zeros(9)
    lib().kfo_jacobian(_d(u), C.byref(g.c()), x, y, z, h, _d(o))
    return o.reshape(3, 3)


def killing_energy(u, g: OGrid, x, y, z, h=0.05, gamma=0.0):
    u = _c64(u)
    return lib().kfo_killing_energy(_d(u), C.byref(g.c()), x, y, z, h, gamma)


def killing_gradient2(u, g: OGrid, x, y, z, h=0.05, gamma=0.0, clamp=0.1):
    u = _c64(u)
    o = np.zeros(3)
    lib().kfo_killing_gradient2(_d(u), C.byref(g.c()), x, y, z, h, gamma, clamp, _d(o))
    return o


def energy_gradient(live, can, u, g: OGrid, p: OParams, x, y, z):
    live, can, u = _c64(live), _c64(can), _c64(u)
    o = np.zeros(3)
    lib().kfo_energy_gradient(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()), x, y, z, _d(o))
    return o


def energy_gradient_grid(live, can, u, g: OGrid, p: OParams, masked=True):
    live, can, u = _c64(live), _c64(can), _c64(u)
    o = np.zeros((3, g.n))
    lib().kfo_energy_gradient_grid(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()),
                                   int(masked), _d(o))
    return o


def _stats(s: Stats) -> dict:
    return dict(iters=s.iters, diverged_index=s.diverged_index, max_update=s.max_update,
                energy=s.energy, active=s.active)


def deform_jacobi(live, can, u, g: OGrid, p: OParams):
    """Returns (u_new, stats)."""
    live, can = _c64(live), _c64(can)
    u = np.array(u, dtype=np.float64, copy=True, order="C")
    st = Stats()
    lib().kfo_deform_jacobi(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()), C.byref(st))
    return u, _stats(st)


def deform_gs_serial(live, can, u, g: OGrid, p: OParams):
    live, can = _c64(live), _c64(can)
    u = np.array(u, dtype=np.float64, copy=True, order="C")
    st = Stats()
    lib().kfo_deform_gs_serial(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()), C.byref(st))
    return u, _stats(st)


def sdf_init(g: OGrid):
    return np.full(g.n, g.oob, dtype=np.float64), np.zeros(g.n, dtype=np.int64)


def integrate(phi, w, g: OGrid, clip, depth, K, C2W, min_depth, max_depth):
    """In-place integration of one depth frame (float32 H x W, metres)."""
    depth = np.ascontiguousarray(depth, dtype=np.float32)
    Km = np.ascontiguousarray(K, dtype=np.float64).reshape(9)
    Cm = np.ascontiguousarray(C2W, dtype=np.float64).reshape(16)
    assert phi.dtype == np.float64 and w.dtype == np.int64
    lib().kfo_integrate(_d(phi), w.ctypes.data_as(_i64p), C.byref(g.c()), clip,
                        depth.ctypes.data_as(C.POINTER(C.c_float)), depth.shape[1], depth.shape[0],
                        _d(Km), _d(Cm), min_depth, max_depth)


def warp(phi, w, u, g: OGrid):
    u = _c64(u)
    assert phi.dtype == np.float64 and w.dtype == np.int64
    lib().kfo_warp(_d(phi), w.ctypes.data_as(_i64p), _d(u), C.byref(g.c()))


def fuse(phi1, w1, phi2, w2, merge=True):
    """SDF::fuse (SDF.cpp:199-241) in place; merge = FUSE_BY_MERGE (config.cpp:7)."""
    phi2 = _c64(phi2)
    w2 = np.ascontiguousarray(w2, dtype=np.int64)
    lib().kfo_fuse_mode(_d(phi1), w1.ctypes.data_as(_i64p), _d(phi2), w2.ctypes.data_as(_i64p), phi1.size,
                        int(bool(merge)))


def fuse_warped(phi1, w1, phi2, w2, u, g: OGrid, merge=True):
    """SDF::fuse(other, disp) (SDF.cpp:244-298) in place; merge = FUSE_BY_MERGE."""
    phi2, u = _c64(phi2), _c64(u)
    w2 = np.ascontiguousarray(w2, dtype=np.int64)
    lib().kfo_fuse_warped_mode(_d(phi1), w1.ctypes.data_as(_i64p), _d(phi2), w2.ctypes.data_as(_i64p),
                               _d(u), C.byref(g.c()), int(bool(merge)))


def marching_cubes(phi, g: OGrid, max_surface, clip, iso=0.0):
    """Returns float64 [T, 3, 3] triangle soup (y negated, reference order)."""
    phi = _c64(phi)
    n = lib().kfo_marching_cubes(_d(phi), C.byref(g.c()), max_surface, clip, iso, None, 0)
    out = np.zeros((max(n, 1), 3, 3))
    lib().kfo_marching_cubes(_d(phi), C.byref(g.c()), max_surface, clip, iso, _d(out), n)
    return out[:n]


def mc_edge_mask(ci):
    return lib().kfo_mc_edge_mask(ci)


def mc_case(ci):
    """triTable[ci] of the oracle's own table (MarchingCubes.h:61-318): 16 edge ids, -1 padded."""
    out = (C.c_int32 * 16)()
    lib().kfo_mc_case(ci, out)
    return list(out)


def adj_energy(live, can, u, g: OGrid, p: OParams, per_voxel=False):
    live, can, u = _c64(live), _c64(can), _c64(u)
    ev = np.zeros(g.n) if per_voxel else None
    act = C.c_int64(0)
    e = lib().kfo_adj_energy(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()),
                             _d(ev) if per_voxel else None, C.byref(act))
    return (e, ev, act.value) if per_voxel else e


def adj_energy_local(live, can, u, g: OGrid, p: OParams, x: int, y: int, z: int):
    """The part of the adjoint energy that depends on u at (x, y, z) (arrays must already be
    contiguous float64)."""
    return lib().kfo_adj_energy_local(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()), x, y, z)


def adj_gradient(live, can, u, g: OGrid, p: OParams):
    live, can, u = _c64(live), _c64(can), _c64(u)
    grad = np.zeros((3, g.n))
    mask = np.zeros(g.n, dtype=np.uint8)
    lib().kfo_adj_gradient(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()), _d(grad),
                           mask.ctypes.data_as(_u8p))
    return grad, mask


def adj_deform(live, can, u, g: OGrid, p: OParams):
    live, can = _c64(live), _c64(can)
    u = np.array(u, dtype=np.float64, copy=True, order="C")
    st = Stats()
    lib().kfo_adj_deform(_d(live), _d(can), _d(u), C.byref(g.c()), C.byref(p.c()), C.byref(st))
    return u, _stats(st)


def num_threads() -> int:
    return lib().kfo_num_threads()


def set_num_threads(n: int) -> None:
    lib().kfo_set_num_threads(int(n))


# ---------------------------------------------------------------------------
# The same oracle evaluated in FP32 (build/libkf_oracle_f32.so from kf_oracle_f32.c): every
# double of kf_oracle.c, structs and arrays included, read as float.  Used only to measure
# what FP32 rounding alone does to the reference optimiser.
# ---------------------------------------------------------------------------

_LIB32_PATH = os.path.join(_HERE, "build", "libkf_oracle_f32.so")


def _float_struct(name, base):
    fields = [(f, C.c_float if t is C.c_double else (C.c_float * t._length_ if getattr(t, "_type_", None) is C.c_double
                                                     else t)) for f, t in base._fields_]
    return type(name, (C.Structure,), {"_fields_": fields})


Grid32 = _float_struct("Grid32", Grid)
Params32 = _float_struct("Params32", Params)
Stats32 = _float_struct("Stats32", Stats)
_lib32 = None


def lib32():
    global _lib32
    if _lib32 is None:
        if not os.path.exists(_LIB32_PATH):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = C.CDLL(_LIB32_PATH)
        fp = C.POINTER(C.c_float)
        L.kfo_deform_jacobi.restype = C.c_int
        L.kfo_deform_jacobi.argtypes = [fp, fp, fp, C.POINTER(Grid32), C.POINTER(Params32), C.POINTER(Stats32)]
        _lib32 = L
    return _lib32


def _to32(struct32, src):
    out = struct32()
    for f, _ in src._fields_:
        v = getattr(src, f)
        if hasattr(v, "__len__"):
            for i in range(len(v)):
                getattr(out, f)[i] = v[i]
        else:
            setattr(out, f, v)
    return out


def deform_jacobi_f32(live, can, u, g: OGrid, p: OParams):
    """kfo_deform_jacobi evaluated in FP32 throughout; returns (u_new as float64, stats)."""
    fp = C.POINTER(C.c_float)
    live = np.ascontiguousarray(live, dtype=np.float32)
    can = np.ascontiguousarray(can, dtype=np.float32)
    u32 = np.array(u, dtype=np.float32, copy=True, order="C")
    st = Stats32()
    lib32().kfo_deform_jacobi(live.ctypes.data_as(fp), can.ctypes.data_as(fp), u32.ctypes.data_as(fp),
                              C.byref(_to32(Grid32, g.c())), C.byref(_to32(Params32, p.c())), C.byref(st))
    return u32.astype(np.float64), dict(iters=st.iters, diverged_index=st.diverged_index,
                                        max_update=float(st.max_update), active=st.active)
