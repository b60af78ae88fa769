"""The integration kernel's shortcuts built for the host (tests/cpp/integrate_check.cpp):

* kf_pixel_fast, the pixel decision without FP64 divisions, and kf_pixel_rowwise, its form for
  cameras whose depth and image row stay fixed along a grid row (the reference's identity pose),
  never claim a pixel other than kf_pixel_exact's (the reference's FP64 projection and float
  roundf, SDF.cpp:158-175), nor the rowwise form a camera depth other than the voxel's own, on
  random cameras and at voxels placed next to pixel boundaries, with the float reciprocal at the
  ends of its error range; and they decide nearly every voxel themselves, and stay off for
  images of 4096 columns or rows and for non-finite camera or grid data;
* enc_r, the BF16/FP16 encoding through a double reciprocal of the scale, equals enc (float
  division) bit for bit over random float bit patterns and SDF-sized values at many scales."""
import os
import subprocess

import pytest

from killingfusion_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def check_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("integ") / "integrate_check")
    cmd = [build.hipcc(), "--offload-host-only", "-x", "hip", "-std=c++17", "-O2", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "killingfusion_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "integrate_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fast_pixel_never_disagrees(check_exe, seed):
    r = subprocess.run([check_exe, "project", str(seed), "120"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    voxels, fast, bad, rvoxels, rfast, rowwise = map(int, r.stdout.split())
    print(f"{voxels} voxel checks ({rowwise} of the rowwise form), {fast} decided by a fast form, "
          f"{bad} mismatches; random voxels: {rfast} of {rvoxels} decided fast", flush=True)
    assert bad == 0, r.stderr
    assert rowwise >= 0.4 * voxels
    # the fast path must decide nearly every voxel of a random camera, or the kernel gains nothing
    # (the boundary-placed voxels are meant to fall back)
    assert rfast >= 0.95 * rvoxels


def test_fast_forms_stay_off_where_unproven(check_exe):
    r = subprocess.run([check_exe, "guards"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cases, off = map(int, r.stdout.split())
    assert cases == off == 6


def test_encode_by_reciprocal_equals_division(check_exe):
    r = subprocess.run([check_exe, "quot", "7", "40000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    values, bad = map(int, r.stdout.split())
    print(f"{values} values, {bad} mismatches", flush=True)
    assert bad == 0, r.stderr
