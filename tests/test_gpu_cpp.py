"""The C++ host mirror (include/kf.hpp) on the GPU: tests/cpp/kf_hpp_check runs the reference-
style self-checks through it (plane integration, fuse(self), zero warp, field arithmetic,
deform, the per-frame driver, OFF and .bin dumps); this reads its dumps back."""
import os
import subprocess

import numpy as np
import pytest

from killingfusion_amd import formats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "kf_hpp_check")


def test_cpp_mirror_self_checks(tmp_path):
    print("case: C++ mirror self-checks (tests/cpp/kf_hpp_check)", flush=True)
    assert os.path.exists(EXE), "build first: make -C tests/cpp"
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.startswith("OK")
    # the plane the program integrated: phi = D - Z_c within the truncation band
    sdf = formats.read_sdf_bin(tmp_path / "plane.sdf.bin")
    assert sdf["dims"] == (32, 24, 40) and sdf["voxel"] == 0.01 and sdf["truncation"] == 3.0
    z = 0.01 * (np.arange(40) + 0.5) + 0.5
    dist = np.float64(np.float32(0.7)) - z
    exp = np.where(dist >= -0.03, dist.astype(np.float32), np.float32(0.03 + 1e-5)).astype(np.float64)
    assert np.array_equal(sdf["phi"].reshape(40, 24, 32), np.broadcast_to(exp[:, None, None], (40, 24, 32)))
    u = formats.read_displacement_bin(tmp_path / "field.bin")
    assert u["dims"] == (32, 24, 40) and np.array_equal(u["u"][:, 0], [0.75, 0.0, 1.5])
    v, f = formats.read_off(tmp_path / "plane.off")
    assert f.shape[0] > 100 and v.shape[0] == 3 * f.shape[0]
    # the driver's per-frame field dump of its last frame equals a dump of its field afterwards
    assert (tmp_path / "frame_014.bin").read_bytes() == (tmp_path / "last_field.bin").read_bytes()
    last = formats.read_displacement_bin(tmp_path / "frame_014.bin")
    assert last["dims"] == (32, 24, 40) and np.isfinite(last["u"]).all()
    assert not (tmp_path / "frame_010.bin").exists()


def test_python_and_cpp_dumps_are_byte_identical(ctx, tmp_path):
    print("case: SDF / field dumps from the Python mirror vs the C++ mirror", flush=True)
    from killingfusion_amd import SDF, DisplacementField, scenes
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    plane = SDF(ctx, (32, 24, 40), (-0.16, -0.12, 0.5), 0.01, 0.03, 0.03)
    plane.integrateDepthFrame(np.full((480, 640), 0.7, dtype=np.float32), scenes.C2W, scenes.K, 0.1, 3.0)
    plane.dumpToBinFile(tmp_path / "py.sdf.bin", 3.0, 0.0)
    assert (tmp_path / "py.sdf.bin").read_bytes() == (tmp_path / "plane.sdf.bin").read_bytes()
    a = DisplacementField(ctx, (32, 24, 40))
    a.initializeAllVoxels((0.75, 0.0, 1.5))
    a.dumpToBinFile(tmp_path / "py.field.bin", 0.01)
    assert (tmp_path / "py.field.bin").read_bytes() == (tmp_path / "field.bin").read_bytes()
