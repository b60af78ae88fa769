"""ctypes bindings of libkf.so (the C ABI declared in include/kf_api.h).

There is no fallback: if the HIP library has not been built, importing the bindings
raises, so nothing can silently run a CPU path in its place.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkf.so")

KF_OK, KF_E_INVALID_ARG, KF_E_DIVERGED, KF_E_OOM, KF_E_HIP, KF_E_NO_DEVICE, KF_E_CAPACITY = range(7)
KF_F32, KF_BF16, KF_F16 = 0, 1, 2
KF_METHOD_REFERENCE, KF_METHOD_ADJOINT = 0, 1
KF_SWEEP_AUTO, KF_SWEEP_DENSE, KF_SWEEP_ACTIVE = 0, 1, 2
KF_MESH_LIVE, KF_MESH_LIVE_WARPED, KF_MESH_CANONICAL = 1, 2, 4
KF_FUSE_MERGE, KF_FUSE_NEAREST = 0, 1

DTYPES = {"f32": KF_F32, "bf16": KF_BF16, "f16": KF_F16}


class VolumeDesc(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("min_corner", C.c_double * 3), ("voxel_size", C.c_double),
                ("clip_distance", C.c_double), ("max_surface", C.c_double), ("epsilon", C.c_double),
                ("dtype", C.c_int32), ("reserved", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("voxel_size", C.c_double), ("max_surface", C.c_double), ("clip_distance", C.c_double),
                ("epsilon", C.c_double), ("alpha", C.c_double), ("omega_killing", C.c_double),
                ("gamma_killing", C.c_double), ("omega_levelset", C.c_double), ("delta_size", C.c_double),
                ("killing_clamp", C.c_double), ("use_data", C.c_int32), ("use_levelset", C.c_int32),
                ("use_killing", C.c_int32), ("max_iters", C.c_int32), ("stop_max_update", C.c_double),
                ("method", C.c_int32), ("sweep", C.c_int32), ("use_trust", C.c_int32), ("reserved", C.c_int32)]


class DeformStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("sweep", C.c_int32), ("diverged_index", C.c_int64),
                ("max_update", C.c_double), ("energy", C.c_double), ("active_voxels", C.c_uint64),
                ("listed_voxels", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("voxels_per_thread", C.c_int32), ("blocks", C.c_int32), ("candidates", C.c_int32),
                ("reserved", C.c_int32), ("us_per_iteration", C.c_double)]


class RigidStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("reserved", C.c_int32), ("voxels", C.c_uint64),
                ("initial_energy", C.c_double), ("energy", C.c_double), ("last_step", C.c_double)]


class PipelineDesc(C.Structure):
    _fields_ = [("volume", VolumeDesc), ("params", Params), ("K", C.c_double * 9), ("C2W", C.c_double * 16),
                ("min_depth", C.c_double), ("max_depth", C.c_double), ("width", C.c_int32),
                ("height", C.c_int32), ("fuse_mode", C.c_int32), ("reserved", C.c_int32)]


class FrameStats(C.Structure):
    _fields_ = [("frame_index", C.c_int32), ("reserved", C.c_int32), ("deform", DeformStats),
                ("ms_integrate", C.c_double), ("ms_deform", C.c_double), ("ms_fuse", C.c_double),
                ("ms_total", C.c_double)]


_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every symbol include/kf_api.h declares.
SIGNATURES = {
    "kf_status_string": (C.c_char_p, [C.c_int]),
    "kf_last_error": (C.c_char_p, []),
    "kf_api_version": (C.c_int, []),
    "kf_params_default": (C.c_int, [C.POINTER(Params)]),
    "kf_context_create": (C.c_int, [C.c_int, _vp, C.POINTER(_vp)]),
    "kf_context_destroy": (C.c_int, [_vp]),
    "kf_context_synchronize": (C.c_int, [_vp]),
    "kf_context_stream": (_vp, [_vp]),
    "kf_context_set_loop_capture": (C.c_int, [_vp, C.c_int32]),
    "kf_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
    "kf_host_free": (C.c_int, [_vp]),
    "kf_dims_from_bounds": (C.c_int, [_dp, _dp, C.c_double, _i32p]),
    "kf_volume_create": (C.c_int, [_vp, C.POINTER(VolumeDesc), C.POINTER(_vp)]),
    "kf_volume_destroy": (C.c_int, [_vp]),
    "kf_volume_describe": (C.c_int, [_vp, C.POINTER(VolumeDesc)]),
    "kf_volume_reset": (C.c_int, [_vp]),
    "kf_volume_upload": (C.c_int, [_vp, _fp, _i32p]),
    "kf_volume_download": (C.c_int, [_vp, _fp, _i32p]),
    "kf_volume_copy": (C.c_int, [_vp, _vp]),
    "kf_field_create": (C.c_int, [_vp, _i32p, C.POINTER(_vp)]),
    "kf_field_destroy": (C.c_int, [_vp]),
    "kf_field_fill": (C.c_int, [_vp, _dp]),
    "kf_field_upload": (C.c_int, [_vp, _fp]),
    "kf_field_download": (C.c_int, [_vp, _fp]),
    "kf_field_add": (C.c_int, [_vp, _vp]),
    "kf_integrate_depth": (C.c_int, [_vp, _fp, C.c_int32, C.c_int32, _dp, _dp, C.c_double, C.c_double]),
    "kf_energy_gradient_ref": (C.c_int, [_vp, _vp, _vp, C.POINTER(Params), _vp]),
    "kf_energy_forward": (C.c_int, [_vp, _vp, _vp, C.POINTER(Params), _dp, _fp]),
    "kf_energy_backward": (C.c_int, [_vp, _vp, _vp, C.POINTER(Params), _vp]),
    "kf_deform": (C.c_int, [_vp, _vp, _vp, C.POINTER(Params), C.POINTER(DeformStats)]),
    "kf_warp": (C.c_int, [_vp, _vp]),
    "kf_fuse": (C.c_int, [_vp, _vp]),
    "kf_fuse_warped": (C.c_int, [_vp, _vp, _vp]),
    "kf_fuse_mode": (C.c_int, [_vp, _vp, C.c_int32]),
    "kf_fuse_warped_mode": (C.c_int, [_vp, _vp, _vp, C.c_int32]),
    "kf_marching_cubes": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_uint64), _fp, C.c_uint64]),
    "kf_marching_cubes_indexed": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _fp, _fp,
                                            C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64]),
    "kf_batch_deform": (C.c_int, [C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.c_int32,
                                  C.POINTER(Params), C.POINTER(DeformStats)]),
    "kf_batch_autotune": (C.c_int, [C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.c_int32,
                                    C.POINTER(Params), C.c_int32, C.POINTER(BatchConfig)]),
    "kf_rigid_align": (C.c_int, [_vp, _vp, C.c_int32, _dp, C.POINTER(RigidStats)]),
    "kf_field_from_rigid": (C.c_int, [_vp, _vp, _vp, _dp]),
    "kf_pipeline_create": (C.c_int, [_vp, C.POINTER(PipelineDesc), C.POINTER(_vp)]),
    "kf_pipeline_destroy": (C.c_int, [_vp]),
    "kf_pipeline_process_frame": (C.c_int, [_vp, _fp, C.POINTER(FrameStats)]),
    "kf_marching_cubes_noalloc": (C.c_int, [_vp, C.c_double, _vp, C.c_uint64, _vp]),
    "kf_pipeline_set_meshes": (C.c_int, [_vp, C.c_int32, C.c_double]),
    "kf_pipeline_mesh": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_uint64), _fp, C.c_uint64]),
    "kf_pipeline_canonical": (_vp, [_vp]),
    "kf_pipeline_live": (_vp, [_vp]),
    "kf_pipeline_field": (_vp, [_vp]),
}

_lib = None


class KFError(RuntimeError):
    def __init__(self, status: int, where: str, message: str = ""):
        self.status = status
        super().__init__(f"{where}: status {status} ({message})")


def load():
    """Loads libkf.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m killingfusion_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(status: int, where: str):
    if status != KF_OK:
        L = load()
        msg = L.kf_last_error().decode(errors="replace")
        raise KFError(status, where, msg)
