"""The reference's on-disk formats (CPU): byte layouts restated from SDF.cpp:432-460 and
DisplacementField.cpp:349-366, and the OFF text of SimpleMesh.h:113-141."""
import struct

import numpy as np

from killingfusion_amd import formats


def test_sdf_bin_layout_and_round_trip(tmp_path):
    dims, mn = (3, 2, 2), (-0.1, 0.2, 0.5)
    phi = np.arange(12, dtype=np.float64) * 0.001 - 0.004
    p = tmp_path / "a.bin"
    formats.write_sdf_bin(p, dims, mn, 0.012, 4.0, phi)
    raw = p.read_bytes()
    assert len(raw) == 12 + 24 + 16 + 8 * 12
    assert struct.unpack("<3i", raw[:12]) == dims
    assert struct.unpack("<3d", raw[12:36]) == mn
    assert struct.unpack("<2d", raw[36:52]) == (0.012, 4.0)
    assert struct.unpack("<12d", raw[52:]) == tuple(phi)
    r = formats.read_sdf_bin(p)
    assert r["dims"] == dims and r["voxel"] == 0.012 and r["truncation"] == 4.0
    assert np.array_equal(r["phi"], phi)


def test_displacement_bin_is_per_voxel_xyz(tmp_path):
    dims = (2, 2, 1)
    u = np.array([[1.0, 2, 3, 4], [10, 20, 30, 40], [100, 200, 300, 400]])
    p = tmp_path / "u.bin"
    formats.write_displacement_bin(p, dims, 0.01, u)
    raw = p.read_bytes()
    assert len(raw) == 12 + 8 + 24 * 4
    assert struct.unpack("<3i", raw[:12]) == dims and struct.unpack("<d", raw[12:20]) == (0.01,)
    assert struct.unpack("<6d", raw[20:68]) == (1.0, 10.0, 100.0, 2.0, 20.0, 200.0)
    r = formats.read_displacement_bin(p)
    assert np.array_equal(r["u"], u) and r["dims"] == dims


def test_off_text_and_round_trip(tmp_path):
    tris = np.array([[[0.0, -0.5, 1.0], [0.1234567, 2e-5, -3.0], [1, 1, 1]]])
    p = tmp_path / "m.off"
    formats.write_off(p, tris)
    lines = p.read_text().splitlines()
    assert lines[:2] == ["OFF", "3 1 0"]
    assert lines[2:5] == ["0 -0.5 1", "0.123457 2e-05 -3", "1 1 1"]  # C++ stream default (%g)
    assert lines[5] == "3 0 1 2"
    v, f = formats.read_off(p)
    assert f.tolist() == [[0, 1, 2]] and np.allclose(v, tris.reshape(3, 3), rtol=5e-6, atol=1e-12)
