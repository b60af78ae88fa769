"""CPU-side checks of the C ABI: libkf.so loads, exports every symbol include/kf_api.h
declares, and rejects bad arguments before touching a device.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from killingfusion_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "kf_api.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kf_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 35
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding table covers them all
    assert set(names) <= set(_lib.SIGNATURES)


def test_params_default_is_reference_config():
    L = _lib.load()
    p = _lib.Params()
    assert L.kf_params_default(C.byref(p)) == _lib.KF_OK
    # config.cpp:4-57
    assert p.voxel_size == 0.012 and p.max_surface == 0.012 * 4 and p.clip_distance == 0.012 * 4
    assert p.alpha == 0.025 and p.omega_killing == 0.04 and p.gamma_killing == 0.0
    assert p.omega_levelset == 0.1 and p.delta_size == 0.05 and p.epsilon == 1e-5
    assert p.max_iters == 1024 and p.killing_clamp == 0.1
    assert (p.use_data, p.use_levelset, p.use_killing) == (1, 1, 1)


def test_dims_from_bounds_matches_reference_rule():
    import oracle as O
    from killingfusion_amd import dims_from_bounds
    # SDF.cpp:122, including the truncation case where bound/voxel lands just below an integer
    for mn, mx, v in [((0, 0, 0), (8, 8, 8), 0.5), ((-0.3, -0.2, 0.5), (0.3, 0.25, 1.2), 0.012),
                      ((0, 0, 0), (0.3, 0.3, 0.3), 0.1)]:
        assert dims_from_bounds(mn, mx, v) == O.dims_from_bounds(mn, mx, v)


def test_status_strings_and_errors_without_device():
    L = _lib.load()
    assert L.kf_status_string(_lib.KF_E_DIVERGED) == b"displacement field diverged"
    assert L.kf_api_version() == 1
    # null pointers are rejected before any device work
    assert L.kf_volume_create(None, None, None) == _lib.KF_E_INVALID_ARG
    assert L.kf_deform(None, None, None, None, None) == _lib.KF_E_INVALID_ARG
    assert L.kf_marching_cubes(None, 0.0, None, None, 0) == _lib.KF_E_INVALID_ARG
    assert L.kf_marching_cubes_indexed(None, 0.0, None, None, None, None, None, 0, 0) == _lib.KF_E_INVALID_ARG
    assert L.kf_integrate_depth(None, None, 640, 480, None, None, 0.1, 3.0) == _lib.KF_E_INVALID_ARG
    assert L.kf_batch_deform(None, None, None, 2, None, None) == _lib.KF_E_INVALID_ARG
    assert L.kf_rigid_align(None, None, 5, None, None) == _lib.KF_E_INVALID_ARG
    assert L.kf_field_from_rigid(None, None, None, None) == _lib.KF_E_INVALID_ARG
    assert L.kf_batch_autotune(None, None, None, 2, None, 0, None) == _lib.KF_E_INVALID_ARG
    bad = np.zeros(3, dtype=np.float64)
    out = (C.c_int32 * 3)()
    assert L.kf_dims_from_bounds(bad.ctypes.data_as(C.POINTER(C.c_double)),
                                 bad.ctypes.data_as(C.POINTER(C.c_double)), 0.0, out) == _lib.KF_E_INVALID_ARG
    assert b"bad bounds" in L.kf_last_error()


def test_context_without_gpu_reports_no_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    L = _lib.load()
    h = C.c_void_p()
    assert L.kf_context_create(0, None, C.byref(h)) == _lib.KF_E_NO_DEVICE
    assert not h.value


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libkf.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.load()


def test_python_mirror_rejects_short_host_buffers():
    # The size checks run before any C call (they must survive python -O, so they are not
    # asserts): a short buffer would otherwise be read past its end by the upload copy.
    from killingfusion_amd import SDF, DisplacementField
    v = SDF.__new__(SDF)
    v.dims, v._L = (3, 4, 5), None
    with pytest.raises(ValueError):
        v.upload(np.zeros(59, dtype=np.float32))
    with pytest.raises(ValueError):
        v.upload(np.zeros(60, dtype=np.float32), np.zeros(59, dtype=np.int32))
    f = DisplacementField.__new__(DisplacementField)
    f.dims, f._L = (3, 4, 5), None
    with pytest.raises(ValueError):
        f.upload(np.zeros((3, 59), dtype=np.float32))
    with pytest.raises(ValueError):
        f.initializeAllVoxels((0.1, 0.2))
    with pytest.raises(ValueError):
        f.initializeAllVoxels(np.zeros(4))
