"""U32Div, the multiply-high division by a divisor fixed before a launch (kf_internal.hpp), built
for the host and compared with exact integer division: every voxel-index decode of the kernels
(GridDims::xyz32, kf_xyz: z = i / (nx ny), y = r / nx) goes through it, so a wrong quotient for
any index would misplace a voxel.  Divisors cover 1, powers of two and their neighbours, every
value up to 2048, grid rows and slices of the configurations, and random 32-bit values;
numerators cover 0, the multiples of d and their neighbours, 2^32 - 1 and random values."""
import os
import subprocess

import numpy as np
import pytest

from killingfusion_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def udiv_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("udiv") / "udiv_check")
    cmd = [build.hipcc(), "--offload-host-only", "-x", "hip", "-std=c++17", "-O2",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "killingfusion_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "udiv_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def divisors():
    rng = np.random.default_rng(5)
    ds = set(range(1, 2049))
    for k in range(32):
        p = 1 << k
        ds.update(v for v in (p - 1, p, p + 1) if 1 <= v < 1 << 32)
    for n in (48, 64, 77, 128, 160, 192, 256, 320, 384, 512, 1024):
        ds.update((n, n * n, n * (n - 1)))
    ds.update(int(v) for v in rng.integers(1, 1 << 32, 2000, dtype=np.uint64))
    ds.update((0xFFFFFFFF, 0xFFFFFFFE, 0x80000001, 0x7FFFFFFF))
    return sorted(ds)


def numerators(d, rng):
    top = (1 << 32) - 1
    ns = {0, 1, top, top - 1, d - 1, d, d + 1}
    for q in (1, 2, 3, top // d, top // d - 1, int(rng.integers(0, top // d + 1))):
        for e in (-1, 0, 1):
            ns.add(q * d + e)
    ns.update(int(v) for v in rng.integers(0, top, 6, dtype=np.uint64, endpoint=True))
    return sorted(n for n in ns if 0 <= n <= top)


def test_u32div_matches_exact_division(udiv_exe):
    rng = np.random.default_rng(11)
    pairs = [(d, n) for d in divisors() for n in numerators(d, rng)]
    inp = "".join(f"{d} {n}\n" for d, n in pairs)
    out = subprocess.run([udiv_exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    assert len(out) == 2 * len(pairs)
    bad = [(d, n, int(out[2 * k]), int(out[2 * k + 1])) for k, (d, n) in enumerate(pairs)
           if int(out[2 * k]) != n // d or int(out[2 * k + 1]) != n % d]
    assert not bad, bad[:10]
    print(f"{len(pairs)} (d, n) pairs exact", flush=True)
