"""The reference's on-disk formats, read and written with numpy (host side).

* SDF grid, SDF::dumpToBinFile (SDF.cpp:432-460): int32 dims[3], float64 min corner[3],
  float64 voxel size, float64 truncation (as passed), then N float64 distances (x fastest).
  Weights are not stored.
* Displacement field, DisplacementField::dumpToBinFile (DisplacementField.cpp:349-366):
  int32 dims[3], float64 voxel size, then N x 3 float64 (x, y, z per voxel).
* Mesh, SimpleMesh::WriteMesh / SimpleMesh(filepath) (SimpleMesh.h:25-58, 113-141): OFF text,
  "OFF", "nv nt 0", one "x y z" line per vertex with 6 significant digits (the default of a
  C++ stream), one "3 i j k" line per face.
All binary data is native (little-endian) byte order, as the reference writes it.
"""
from __future__ import annotations

import numpy as np


def write_sdf_bin(path, dims, min_corner, voxel, truncation, phi):
    dims = np.asarray(dims, dtype=np.int32)
    phi = np.asarray(phi, dtype=np.float64).reshape(-1)
    if phi.size != int(np.prod(dims, dtype=np.int64)):
        raise ValueError("phi must hold dims[0] * dims[1] * dims[2] values")
    with open(path, "wb") as f:
        f.write(dims.tobytes())
        f.write(np.asarray(min_corner, dtype=np.float64).tobytes())
        f.write(np.asarray([voxel, truncation], dtype=np.float64).tobytes())
        f.write(phi.tobytes())


def read_sdf_bin(path) -> dict:
    with open(path, "rb") as f:
        dims = np.frombuffer(f.read(12), dtype=np.int32).copy()
        mn = np.frombuffer(f.read(24), dtype=np.float64).copy()
        voxel, trunc = np.frombuffer(f.read(16), dtype=np.float64)
        n = int(np.prod(dims, dtype=np.int64))
        phi = np.frombuffer(f.read(8 * n), dtype=np.float64)
        if phi.size != n or f.read(1):
            raise ValueError(f"{path}: size does not match dims {tuple(dims)}")
    return dict(dims=tuple(int(d) for d in dims), min_corner=tuple(mn), voxel=float(voxel),
                truncation=float(trunc), phi=phi.copy())


def write_displacement_bin(path, dims, voxel, u_soa):
    dims = np.asarray(dims, dtype=np.int32)
    n = int(np.prod(dims, dtype=np.int64))
    u = np.asarray(u_soa, dtype=np.float64).reshape(3, n)
    with open(path, "wb") as f:
        f.write(dims.tobytes())
        f.write(np.asarray([voxel], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(u.T).tobytes())


def read_displacement_bin(path) -> dict:
    """Returns dims, voxel and u as SoA float64 [3, N]."""
    with open(path, "rb") as f:
        dims = np.frombuffer(f.read(12), dtype=np.int32).copy()
        voxel = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        n = int(np.prod(dims, dtype=np.int64))
        aos = np.frombuffer(f.read(24 * n), dtype=np.float64)
        if aos.size != 3 * n or f.read(1):
            raise ValueError(f"{path}: size does not match dims {tuple(dims)}")
    return dict(dims=tuple(int(d) for d in dims), voxel=voxel, u=aos.reshape(n, 3).T.copy())


def write_off(path, tris):
    """tris: [T, 3, 3] triangle soup; vertex 3t + k is corner k of triangle t."""
    tris = np.asarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    t = tris.shape[0]
    lines = ["OFF", f"{3 * t} {t} 0"]
    lines += ["%g %g %g" % tuple(v) for v in tris.reshape(-1, 3)]
    lines += [f"3 {3 * i} {3 * i + 1} {3 * i + 2}" for i in range(t)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_off_indexed(path, vertices, faces):
    """OFF of an indexed mesh: vertices [V, 3], faces [F, 3]."""
    v = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    lines = ["OFF", f"{v.shape[0]} {f.shape[0]} 0"]
    lines += ["%g %g %g" % tuple(x) for x in v]
    lines += ["3 %d %d %d" % tuple(t) for t in f]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_off(path):
    """Returns (vertices [V, 3] float64, faces [F, 3] int64)."""
    with open(path) as f:
        tok = f.read().split()
    if not tok or tok[0] != "OFF":
        raise ValueError(f"{path}: not an OFF file")
    nv, nf = int(tok[1]), int(tok[2])
    p = 4
    v = np.array(tok[p:p + 3 * nv], dtype=np.float64).reshape(nv, 3)
    p += 3 * nv
    fc = np.array(tok[p:p + 4 * nf], dtype=np.int64).reshape(nf, 4)
    if (fc[:, 0] != 3).any() or (fc[:, 1:] >= nv).any() or (fc[:, 1:] < 0).any():
        raise ValueError(f"{path}: only triangles with valid indices are supported")
    return v, fc[:, 1:]
