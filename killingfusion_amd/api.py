"""Python mirror of the reference's SDF / DisplacementField / KillingFusion interface over
libkf's C ABI (include/kf_api.h).

Class and method names follow the reference (KillingFusion/include/SDF.h:16-200,
DisplacementField.h:9-66, KillingFusion.h:13-100) so tests and callers read like the
reference code; every call runs the HIP kernels of libkf.so on the context's GPU.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, fields

import numpy as np

from . import _lib, formats
from ._lib import check

_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


_SWEEPS = {"auto": 0, "dense": 1, "active": 2}  # kf_sweep


def host_array(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in page-locked host memory (kf_host_alloc), freed with the array: frames
    kept in it upload to the GPU by DMA, without a staging copy."""
    L = _lib.load()
    dt = np.dtype(dtype)
    n = int(np.prod(shape)) * dt.itemsize
    ptr = C.c_void_p()
    check(L.kf_host_alloc(max(n, 1), C.byref(ptr)), "kf_host_alloc")
    buf = (C.c_char * max(n, 1)).from_address(ptr.value)
    arr = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)
    weakref.finalize(buf, L.kf_host_free, C.c_void_p(ptr.value))
    return arr


@dataclass
class KillingParams:
    """Optimiser parameters (the reference's config.cpp:4-57 globals as a runtime struct).
    Defaults are the reference's; `paper()` gives the values quoted in config.cpp:32-41."""
    voxel_size: float = 0.012
    max_surface: float = 0.048
    clip_distance: float = 0.048
    epsilon: float = 1e-5
    alpha: float = 0.025
    omega_killing: float = 0.04
    gamma_killing: float = 0.0
    omega_levelset: float = 0.1
    delta_size: float = 0.05
    killing_clamp: float = 0.1
    use_data: bool = True
    use_levelset: bool = True
    use_killing: bool = True
    max_iters: int = 1024
    stop_max_update: float = 0.0
    method: str = "reference"  # or "adjoint"
    sweep: str = "auto"  # "auto", "dense" (every voxel) or "active" (the voxels that can move): the same field
    use_trust: bool = False  # UseTrustStrategy (config.cpp:27): reference method, data energy alone

    @classmethod
    def paper(cls, voxel_size: float, trunc_voxels: float = 3.0, **kw):
        base = dict(voxel_size=voxel_size, max_surface=trunc_voxels * voxel_size,
                    clip_distance=trunc_voxels * voxel_size, alpha=0.1, omega_killing=0.5,
                    gamma_killing=0.1, omega_levelset=0.2)
        base.update(kw)
        return cls(**base)

    def c(self) -> _lib.Params:
        p = _lib.Params()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "method":
                v = _lib.KF_METHOD_ADJOINT if v == "adjoint" else _lib.KF_METHOD_REFERENCE
            elif f.name == "sweep":
                if v not in _SWEEPS:
                    raise ValueError(f"sweep must be one of {sorted(_SWEEPS)}, not {v!r}")
                v = _SWEEPS[v]
            setattr(p, f.name, v)
        return p


class Context:
    """A device + HIP stream (kf_context_create)."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = C.c_void_p()
        check(self._L.kf_context_create(device, None, C.byref(h)), "kf_context_create")
        self.handle = h
        self.device = device
        self._children = weakref.WeakSet()  # handles to release before the context

    def _adopt(self, obj):
        self._children.add(obj)

    def synchronize(self):
        check(self._L.kf_context_synchronize(self.handle), "kf_context_synchronize")

    def set_loop_capture(self, enable: bool):
        """Replay the active-set iteration loops as captured hipGraphs (kf_context_set_loop_capture;
        off by default: direct launches give the same results and measured as fast or faster)."""
        check(self._L.kf_context_set_loop_capture(self.handle, 1 if enable else 0), "kf_context_set_loop_capture")

    def close(self):
        if self.handle:
            for ch in list(self._children):
                ch.close()
            self._L.kf_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dims_from_bounds(min3d, max3d, voxel):
    """SDF::computeVoxelGridSize (SDF.cpp:117-124)."""
    L = _lib.load()
    a = np.array(min3d, dtype=np.float64)
    b = np.array(max3d, dtype=np.float64)
    out = (C.c_int32 * 3)()
    check(L.kf_dims_from_bounds(a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), voxel, out), "kf_dims_from_bounds")
    return tuple(out)


class SimpleMesh:
    """Triangle soup as SimpleMesh (SimpleMesh.h:20-146): vertices (3T, 3), faces (T, 3)."""

    def __init__(self, tris: np.ndarray):
        self.tris = tris  # (T, 3, 3) float32, y already negated

    @property
    def vertices(self):
        return self.tris.reshape(-1, 3)

    @property
    def triangles(self):
        t = self.tris.shape[0]
        return np.arange(3 * t, dtype=np.uint32).reshape(t, 3)

    def WriteMesh(self, filename) -> bool:
        """SimpleMesh::WriteMesh (SimpleMesh.h:113-141): OFF text."""
        try:
            formats.write_off(filename, self.tris)
        except OSError:
            return False
        return True

    @classmethod
    def from_off(cls, filename) -> "SimpleMesh":
        """SimpleMesh(std::string filepath) (SimpleMesh.h:25-58)."""
        v, f = formats.read_off(filename)
        return cls(v[f].astype(np.float32))


class IndexedMesh:
    """Welded marching-cubes output (kf_marching_cubes_indexed): vertices (V, 3), unit normals
    (V, 3) and triangles (T, 3) as vertex indices, in the soup's triangle order."""

    def __init__(self, vertices: np.ndarray, normals: np.ndarray, triangles: np.ndarray):
        self.vertices, self.normals, self.triangles = vertices, normals, triangles

    def WriteMesh(self, filename) -> bool:
        try:
            formats.write_off_indexed(filename, self.vertices, self.triangles)
        except OSError:
            return False
        return True


class SDF:
    """A TSDF volume on the GPU (SDF.h:16-200)."""

    def __init__(self, ctx: Context, dims, min3d=(0.0, 0.0, 0.0), voxel_size=0.012, clip_distance=0.048,
                 max_surface=0.048, epsilon=1e-5, dtype="f32"):
        self._L = _lib.load()
        self.ctx = ctx
        d = _lib.VolumeDesc()
        for a in range(3):
            d.dims[a] = int(dims[a])
            d.min_corner[a] = float(min3d[a])
        d.voxel_size = voxel_size
        d.clip_distance = clip_distance
        d.max_surface = max_surface
        d.epsilon = epsilon
        d.dtype = _lib.DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        h = C.c_void_p()
        check(self._L.kf_volume_create(ctx.handle, C.byref(d), C.byref(h)), "kf_volume_create")
        self.handle = h
        ctx._adopt(self)
        self.dims = tuple(int(x) for x in dims)
        self.min3d = tuple(float(x) for x in min3d)
        self.voxel_size = voxel_size
        self.clip_distance = clip_distance
        self.max_surface = max_surface
        self.epsilon = epsilon
        self.dtype = dtype

    @classmethod
    def from_bounds(cls, ctx, voxel_size, min3d, max3d, clip_distance, **kw):
        """SDF::SDF(voxelSize, min, max, unknownClipDistance), SDF.cpp:14-30."""
        return cls(ctx, dims_from_bounds(min3d, max3d, voxel_size), min3d, voxel_size, clip_distance, **kw)

    @property
    def n(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    def dumpToBinFile(self, path, truncationDistanceInVoxelSizeUnit: float, minimumWeightThreshold: float = 0.0):
        """SDF::dumpToBinFile (SDF.cpp:432-460); the weight threshold is unused there too."""
        phi, _ = self.download()
        formats.write_sdf_bin(path, self.dims, self.min3d, self.voxel_size, truncationDistanceInVoxelSizeUnit, phi)

    def getGridSize(self):
        return self.dims

    def reset(self):
        check(self._L.kf_volume_reset(self.handle), "kf_volume_reset")

    def upload(self, phi, weight=None):
        phi = _f32(phi).reshape(-1)
        if phi.size != self.n:
            raise ValueError(f"phi holds {phi.size} values, the volume {self.n}")
        wp = None
        if weight is not None:
            w = np.ascontiguousarray(weight, dtype=np.int32).reshape(-1)
            if w.size != self.n:
                raise ValueError(f"weight holds {w.size} values, the volume {self.n}")
            wp = w.ctypes.data_as(_i32p)
        check(self._L.kf_volume_upload(self.handle, phi.ctypes.data_as(_fp), wp), "kf_volume_upload")

    def download(self):
        phi = np.empty(self.n, dtype=np.float32)
        w = np.empty(self.n, dtype=np.int32)
        check(self._L.kf_volume_download(self.handle, phi.ctypes.data_as(_fp), w.ctypes.data_as(_i32p)),
              "kf_volume_download")
        return phi, w

    def copy_from(self, other: "SDF"):
        check(self._L.kf_volume_copy(self.handle, other.handle), "kf_volume_copy")

    def integrateDepthFrame(self, depth, C2W, K, min_depth, max_depth):
        """SDF::integrateDepthFrame (SDF.cpp:137-197); depth float32 (H, W) metres."""
        depth = _f32(depth)
        h, w = depth.shape
        Km = np.ascontiguousarray(K, dtype=np.float64).reshape(9)
        Cm = np.ascontiguousarray(C2W, dtype=np.float64).reshape(16)
        check(self._L.kf_integrate_depth(self.handle, depth.ctypes.data_as(_fp), w, h, Km.ctypes.data_as(_dp),
                                         Cm.ctypes.data_as(_dp), min_depth, max_depth), "kf_integrate_depth")

    def fuse(self, other: "SDF", displacement_field: "DisplacementField" = None, merge: bool = True):
        """SDF::fuse(other) (SDF.cpp:199-241) or SDF::fuse(other, disp) (SDF.cpp:244-298); merge is
        FUSE_BY_MERGE (config.cpp:7): False keeps, per voxel, the value nearer the surface."""
        mode = _lib.KF_FUSE_MERGE if merge else _lib.KF_FUSE_NEAREST
        if displacement_field is None:
            check(self._L.kf_fuse_mode(self.handle, other.handle, mode), "kf_fuse_mode")
        else:
            check(self._L.kf_fuse_warped_mode(self.handle, other.handle, displacement_field.handle, mode),
                  "kf_fuse_warped_mode")

    def update(self, displacement_field: "DisplacementField"):
        """SDF::update (SDF.cpp:300-328): warp by the field."""
        check(self._L.kf_warp(self.handle, displacement_field.handle), "kf_warp")

    def getMesh(self, iso: float = 0.0) -> SimpleMesh:
        """SDF::getMesh (SDF.cpp:396-414)."""
        n = C.c_uint64(0)
        check(self._L.kf_marching_cubes(self.handle, iso, C.byref(n), None, 0), "kf_marching_cubes")
        out = np.empty((max(n.value, 1), 3, 3), dtype=np.float32)
        if n.value:
            check(self._L.kf_marching_cubes(self.handle, iso, C.byref(n), out.ctypes.data_as(_fp), n.value),
                  "kf_marching_cubes")
        return SimpleMesh(out[: n.value])

    def getMeshIndexed(self, iso: float = 0.0, normals: bool = True) -> IndexedMesh:
        """The welded mesh of getMesh (one vertex per cut grid edge) with per-vertex normals."""
        nv, nt = C.c_uint64(0), C.c_uint64(0)
        check(self._L.kf_marching_cubes_indexed(self.handle, iso, C.byref(nv), C.byref(nt), None, None, None, 0, 0),
              "kf_marching_cubes_indexed")
        v = np.empty((max(nv.value, 1), 3), dtype=np.float32)
        n = np.empty((max(nv.value, 1), 3), dtype=np.float32) if normals else None
        t = np.empty((max(nt.value, 1), 3), dtype=np.uint32)
        if nt.value:
            check(self._L.kf_marching_cubes_indexed(
                self.handle, iso, C.byref(nv), C.byref(nt), v.ctypes.data_as(_fp),
                n.ctypes.data_as(_fp) if normals else None, t.ctypes.data_as(C.POINTER(C.c_uint32)), nv.value,
                nt.value), "kf_marching_cubes_indexed")
        return IndexedMesh(v[: nv.value], n[: nv.value] if normals else None, t[: nt.value])

    def close(self):
        if getattr(self, "handle", None):
            self._L.kf_volume_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DisplacementField:
    """Displacement grid in voxel units (DisplacementField.h:9-66), SoA float32 on the GPU."""

    def __init__(self, ctx: Context, dims):
        self._L = _lib.load()
        self.ctx = ctx
        d = (C.c_int32 * 3)(*[int(x) for x in dims])
        h = C.c_void_p()
        check(self._L.kf_field_create(ctx.handle, d, C.byref(h)), "kf_field_create")
        self.handle = h
        ctx._adopt(self)
        self.dims = tuple(int(x) for x in dims)

    @property
    def n(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    def dumpToBinFile(self, path, voxel_size: float = 0.0):
        """DisplacementField::dumpToBinFile (DisplacementField.cpp:349-366)."""
        formats.write_displacement_bin(path, self.dims, voxel_size, self.download())

    def initializeAllVoxels(self, value):
        """DisplacementField::initializeAllVoxels (DisplacementField.cpp:82-88): every voxel = value."""
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
        if v.size != 3:
            raise ValueError(f"a displacement has 3 components, not {v.size}")
        check(self._L.kf_field_fill(self.handle, v.ctypes.data_as(_dp)), "kf_field_fill")

    def upload(self, soa):
        soa = _f32(soa)
        if soa.size != 3 * self.n:
            raise ValueError(f"field upload holds {soa.size} values, the field 3 x {self.n}")
        soa = soa.reshape(3, -1)
        check(self._L.kf_field_upload(self.handle, soa.ctypes.data_as(_fp)), "kf_field_upload")

    def download(self):
        out = np.empty((3, self.n), dtype=np.float32)
        check(self._L.kf_field_download(self.handle, out.ctypes.data_as(_fp)), "kf_field_download")
        return out

    def __iadd__(self, other: "DisplacementField"):
        """DisplacementField::operator+ (DisplacementField.cpp:73-80) adds in place."""
        check(self._L.kf_field_add(self.handle, other.handle), "kf_field_add")
        return self

    def close(self):
        if getattr(self, "handle", None):
            self._L.kf_field_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stats(s: _lib.DeformStats) -> dict:
    return dict(iters=s.iters, diverged_index=s.diverged_index, max_update=s.max_update, energy=s.energy,
                active=s.active_voxels, listed=s.listed_voxels,
                sweep={1: "dense", 2: "active"}.get(s.sweep, "dense"))


def compute_displacement_field(src: SDF, dest: SDF, field: DisplacementField, params: KillingParams) -> dict:
    """KillingFusion::computeDisplacementField (KillingFusion.cpp:240-397), all-voxels mode."""
    L = _lib.load()
    st = _lib.DeformStats()
    p = params.c()
    status = L.kf_deform(src.handle, dest.handle, field.handle, C.byref(p), C.byref(st))
    out = _stats(st)
    if status == _lib.KF_E_DIVERGED:
        raise _lib.KFError(status, "kf_deform", f"diverged at voxel {st.diverged_index}")
    check(status, "kf_deform")
    return out


def energy_gradient_ref(src: SDF, dest: SDF, field: DisplacementField, params: KillingParams,
                        out: DisplacementField):
    """KillingFusion::computeEnergyGradient at every voxel (zero where skipped)."""
    L = _lib.load()
    p = params.c()
    check(L.kf_energy_gradient_ref(src.handle, dest.handle, field.handle, C.byref(p), out.handle),
          "kf_energy_gradient_ref")


def energy_forward(src: SDF, dest: SDF, field: DisplacementField, params: KillingParams, per_voxel=False):
    L = _lib.load()
    p = params.c()
    e = C.c_double(0.0)
    ev = np.empty(src.n, dtype=np.float32) if per_voxel else None
    check(L.kf_energy_forward(src.handle, dest.handle, field.handle, C.byref(p), C.byref(e),
                              ev.ctypes.data_as(_fp) if per_voxel else None), "kf_energy_forward")
    return (e.value, ev) if per_voxel else e.value


def energy_backward(src: SDF, dest: SDF, field: DisplacementField, params: KillingParams,
                    out: DisplacementField):
    L = _lib.load()
    p = params.c()
    check(L.kf_energy_backward(src.handle, dest.handle, field.handle, C.byref(p), out.handle),
          "kf_energy_backward")


def rigid_align(live: SDF, can: SDF, max_iters: int = 30, pose=None):
    """kf_rigid_align: the rigid pose (4x4, canonical world -> live world) that best maps the
    canonical SDF onto the live one, by Gauss-Newton from `pose` (identity by default)."""
    L = _lib.load()
    T = np.ascontiguousarray(np.eye(4) if pose is None else pose, dtype=np.float64).reshape(16).copy()
    st = _lib.RigidStats()
    check(L.kf_rigid_align(live.handle, can.handle, int(max_iters), T.ctypes.data_as(_dp), C.byref(st)),
          "kf_rigid_align")
    return T.reshape(4, 4), dict(iters=st.iters, voxels=st.voxels, initial_energy=st.initial_energy,
                                 energy=st.energy, last_step=st.last_step)


def field_from_rigid(field: DisplacementField, can: SDF, live: SDF, pose):
    """kf_field_from_rigid: the displacement field (live index units) of a rigid pose."""
    L = _lib.load()
    T = np.ascontiguousarray(pose, dtype=np.float64).reshape(16)
    check(L.kf_field_from_rigid(field.handle, can.handle, live.handle, T.ctypes.data_as(_dp)), "kf_field_from_rigid")


def _batch_args(lives, cans, fields_):
    n = len(lives)
    if not (len(cans) == n and len(fields_) == n):
        raise ValueError("batch lists differ in length")
    a = (C.c_void_p * max(n, 1))(*[v.handle.value for v in lives])
    b = (C.c_void_p * max(n, 1))(*[v.handle.value for v in cans])
    c = (C.c_void_p * max(n, 1))(*[f.handle.value for f in fields_])
    return a, b, c, n


def batch_deform(lives, cans, fields_, params: KillingParams, stats: bool = True):
    """kf_batch_deform: compute_displacement_field on independent volumes in one batch.
    Returns one stats dict per volume (None when stats=False, which also skips the final
    energy evaluation)."""
    L = _lib.load()
    a, b, c, n = _batch_args(lives, cans, fields_)
    st = (_lib.DeformStats * max(n, 1))() if stats else None
    p = params.c()
    check(L.kf_batch_deform(a, b, c, n, C.byref(p), st), "kf_batch_deform")
    return [_stats(st[i]) for i in range(n)] if stats else None


def batch_autotune(lives, cans, fields_, params: KillingParams, force: bool = False, shape=None):
    """kf_batch_autotune: times the batched launch's tile shapes on this batch and returns the
    chosen one; shape=(voxels_per_thread, blocks) pins that shape instead (blocks 0 = one
    block per tile)."""
    L = _lib.load()
    a, b, c, n = _batch_args(lives, cans, fields_)
    cfg = _lib.BatchConfig()
    mode = int(bool(force))
    if shape is not None:
        cfg.voxels_per_thread, cfg.blocks = int(shape[0]), int(shape[1])
        mode = 2
    p = params.c()
    check(L.kf_batch_autotune(a, b, c, n, C.byref(p), mode, C.byref(cfg)), "kf_batch_autotune")
    return {"voxels_per_thread": cfg.voxels_per_thread, "blocks": cfg.blocks, "candidates": cfg.candidates,
            "us_per_iteration": cfg.us_per_iteration}


class KillingFusion:
    """Per-frame driver (KillingFusion::processNextFrame, KillingFusion.cpp:109-175) on the GPU."""

    def __init__(self, ctx: Context, dims, min3d, voxel_size, params: KillingParams, K, C2W=None,
                 min_depth=0.1, max_depth=3.0, width=640, height=480, dtype="f32", epsilon=1e-5,
                 fuse_by_merge: bool = True):
        self._L = _lib.load()
        self.ctx = ctx
        d = _lib.PipelineDesc()
        for a in range(3):
            d.volume.dims[a] = int(dims[a])
            d.volume.min_corner[a] = float(min3d[a])
        d.volume.voxel_size = voxel_size
        d.volume.clip_distance = params.clip_distance
        d.volume.max_surface = params.max_surface
        d.volume.epsilon = epsilon
        d.volume.dtype = _lib.DTYPES[dtype]
        d.params = params.c()
        Km = np.asarray(K, dtype=np.float64).reshape(9)
        Cm = np.eye(4).reshape(16) if C2W is None else np.asarray(C2W, dtype=np.float64).reshape(16)
        for i in range(9):
            d.K[i] = Km[i]
        for i in range(16):
            d.C2W[i] = Cm[i]
        d.min_depth, d.max_depth, d.width, d.height = min_depth, max_depth, width, height
        d.fuse_mode = _lib.KF_FUSE_MERGE if fuse_by_merge else _lib.KF_FUSE_NEAREST
        h = C.c_void_p()
        check(self._L.kf_pipeline_create(ctx.handle, C.byref(d), C.byref(h)), "kf_pipeline_create")
        self.handle = h
        ctx._adopt(self)
        self.dims = tuple(dims)
        self.voxel_size = voxel_size
        self._deforms = bool(params.use_data or params.use_levelset or params.use_killing)
        self._dump = None

    def setFieldDumpPrefix(self, prefix, start_frame: int = 0, stride: int = 1):
        """The reference's per-frame displacement dump (KillingFusion.cpp:146-151): after the
        deform of each frame but the first (when some energy is on), the frame's field goes to
        prefix + the dataset frame number (start_frame + stride x the frame's index, 3 digits) +
        ".bin" in DisplacementField::dumpToBinFile's layout.  None (the default) writes nothing."""
        self._dump = None if prefix is None else (str(prefix), int(start_frame), int(stride))

    MESHES = {"live": _lib.KF_MESH_LIVE, "live_warped": _lib.KF_MESH_LIVE_WARPED,
              "canonical": _lib.KF_MESH_CANONICAL}

    def setMeshes(self, which=("live", "live_warped", "canonical"), iso: float = 0.0):
        """Selects the per-frame meshes processNextFrame extracts (kf_pipeline_set_meshes): any of
        "live" (the live SDF before the warp, KillingFusion.cpp:157), "live_warped" (after it,
        :160) and "canonical" (after the fuse, :170)."""
        mask = 0
        for w in which:
            mask |= self.MESHES[w]
        check(self._L.kf_pipeline_set_meshes(self.handle, mask, iso), "kf_pipeline_set_meshes")

    def frameMesh(self, which: str) -> SimpleMesh:
        """One mesh of the last processed frame (kf_pipeline_mesh)."""
        bit = self.MESHES[which]
        n = C.c_uint64(0)
        check(self._L.kf_pipeline_mesh(self.handle, bit, C.byref(n), None, 0), "kf_pipeline_mesh")
        out = np.empty((max(n.value, 1), 3, 3), dtype=np.float32)
        if n.value:
            check(self._L.kf_pipeline_mesh(self.handle, bit, C.byref(n), out.ctypes.data_as(_fp), n.value),
                  "kf_pipeline_mesh")
        return SimpleMesh(out[: n.value])

    def frameMeshes(self):
        """The reference's return value of processNextFrame (KillingFusion.cpp:171-173): the live
        mesh, the warped live mesh and the canonical mesh of the last frame (all three selected)."""
        return [self.frameMesh(w) for w in ("live", "live_warped", "canonical")]

    def processNextFrame(self, depth) -> dict:
        depth = _f32(depth)
        st = _lib.FrameStats()
        check(self._L.kf_pipeline_process_frame(self.handle, depth.ctypes.data_as(_fp), C.byref(st)),
              "kf_pipeline_process_frame")
        if self._dump is not None and st.frame_index > 0 and self._deforms:
            prefix, start, stride = self._dump
            formats.write_displacement_bin(f"{prefix}{start + stride * st.frame_index:03d}.bin", self.dims,
                                           self.voxel_size, self.field_download())
        return dict(frame=st.frame_index, deform=_stats(st.deform), ms_integrate=st.ms_integrate,
                    ms_deform=st.ms_deform, ms_fuse=st.ms_fuse, ms_total=st.ms_total)

    def _volume_download(self, h):
        n = self.dims[0] * self.dims[1] * self.dims[2]
        phi = np.empty(n, dtype=np.float32)
        w = np.empty(n, dtype=np.int32)
        check(self._L.kf_volume_download(h, phi.ctypes.data_as(_fp), w.ctypes.data_as(_i32p)), "kf_volume_download")
        return phi, w

    def canonical_download(self):
        """The canonical SDF (m_canonicalSdf, KillingFusion.h:99): phi (metres), weights."""
        return self._volume_download(self._L.kf_pipeline_canonical(self.handle))

    def live_download(self):
        """The last frame's live SDF, warped into the canonical frame (SDF::update, KillingFusion.cpp:158)."""
        return self._volume_download(self._L.kf_pipeline_live(self.handle))

    def field_download(self):
        """The last frame's displacement field (live -> canonical), SoA float32 [3, N], voxels."""
        n = self.dims[0] * self.dims[1] * self.dims[2]
        out = np.empty((3, n), dtype=np.float32)
        check(self._L.kf_field_download(self._L.kf_pipeline_field(self.handle), out.ctypes.data_as(_fp)),
              "kf_field_download")
        return out

    def canonical_mesh(self, iso=0.0) -> SimpleMesh:
        h = C.c_void_p(self._L.kf_pipeline_canonical(self.handle))
        n = C.c_uint64(0)
        check(self._L.kf_marching_cubes(h, iso, C.byref(n), None, 0), "kf_marching_cubes")
        out = np.empty((max(n.value, 1), 3, 3), dtype=np.float32)
        if n.value:
            check(self._L.kf_marching_cubes(h, iso, C.byref(n), out.ctypes.data_as(_fp), n.value),
                  "kf_marching_cubes")
        return SimpleMesh(out[: n.value])

    def close(self):
        if getattr(self, "handle", None):
            self._L.kf_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
