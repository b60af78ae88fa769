"""kf_split, the cell split of every trilinear read in the kernels (kf_internal.hpp), built for
the host and checked against the reference's read convention: the point i + r reads cell
base = (int)(i + r), truncated toward zero, with weight (i + r) - base (SDF.cpp:503-524,
DisplacementField.cpp:40-58).  Beyond |r| = 2^30 the reference's cast is undefined; there the
cell index is clamped so that every tap lies outside any grid (at most 2^15 per axis) and the
weight keeps the exact floor."""
import math
import os
import subprocess

import numpy as np
import pytest

from killingfusion_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def split_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("split") / "split_check")
    cmd = [build.hipcc(), "--offload-host-only", "-x", "hip", "-std=c++17", "-O2",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "killingfusion_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "split_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def split(exe, pairs):
    inp = "".join(f"{i} {np.float32(r).view(np.uint32):08x}\n" for i, r in pairs)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    res = []
    for line in out:
        if line:
            b, fb = line.split()
            res.append((int(b), np.uint32(int(fb, 16)).view(np.float32)))
    return res


def test_split_follows_truncation(split_exe):
    rng = np.random.default_rng(3)
    i = rng.integers(0, 300, 4000)
    r = np.concatenate([rng.uniform(-3.0, 3.0, 3000), rng.integers(-3, 4, 500),
                        rng.uniform(-1.0, 1.0, 500) * 1e-6]).astype(np.float32)
    got = split(split_exe, list(zip(i.tolist(), r.tolist())))
    for (ii, rr), (b, f) in zip(zip(i.tolist(), r.tolist()), got):
        t = ii + float(np.float32(rr))  # exact in double
        assert b == math.trunc(t) or (t == int(t) and b == int(t)), (ii, rr, b)
        assert float(b) + float(f) == pytest.approx(t, abs=2 ** -22 * max(1.0, abs(rr)))
        # the weight is r - floor(r) rounded to float32, so it may round up to 1 (r just below 0)
        assert (0.0 <= f <= 1.0) if t >= 0 else (-1.0 <= f <= 0.0)


def test_split_clamps_huge_displacements(split_exe):
    pairs = [(0, 3.0e9), (7, 3.0e9), (0, -3.0e9), (200, -3.0e9), (0, 2.0 ** 30 + 128), (5, 1.5e38), (5, -1.5e38)]
    got = split(split_exe, pairs)
    for (i, r), (b, f) in zip(pairs, got):
        assert abs(b - i) == 2 ** 30, (i, r, b)
        assert f == 0.0  # floats this large are integers
    (b, f), = split(split_exe, [(3, float("nan"))])
    assert math.isnan(f) and abs(b - 3) == 2 ** 30
    # below the clamp nothing changes
    (b, f), = split(split_exe, [(4, 2.0 ** 22 + 0.5)])
    assert b == 4 + 2 ** 22 and f == 0.5
    (b, f), = split(split_exe, [(4, 2.0 ** 29)])
    assert b == 4 + 2 ** 29 and f == 0.0
