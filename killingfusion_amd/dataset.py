"""Real-data front end: the reference's DatasetReader and camera-frustum bounds.

* DatasetReader (KillingFusion/src/DatasetReader.cpp:15-67, 122-134): the Slavcheva deformable
  sequences laid out as <root>/<Duck|Snoopy>/{color,depth,omask}_NNNNNN.png with
  <root>/intrinsics_kinect1.txt (9 numbers, row-major K).  Depth PNGs hold 16-bit millimetres,
  converted to metres with a factor 1/1000; the object mask zeroes every pixel where it is 0
  (cv::Mat::copyTo with a mask into a fresh matrix).  Frames are 640 x 480; the depth range
  is the per-dataset table of config.cpp:16-19.
* compute_bounds (KillingFusion::computeBounds, KillingFusion.cpp:478-519): the axis-aligned
  box of the camera frustum between min and max depth, with the camera at the origin.
* pipeline_for: the reference's KillingFusion(DatasetReader) set-up (KillingFusion.cpp:10-31,
  209-231) -- one canonical grid over that box at the configured voxel size -- as a GPU
  per-frame driver.

PNG decoding uses Pillow, which this module needs only when it reads images.
"""
from __future__ import annotations

import os

import numpy as np

from . import api

# config.cpp:13-19: image count and depth range of each dataset.
DATASETS = {
    "Duck": dict(frames=447, min_depth=0.0497470, max_depth=1.2),
    "Snoopy": dict(frames=630, min_depth=0.0495164, max_depth=1.2),
}
WIDTH, HEIGHT = 640, 480
DEPTH_SCALE = 1.0 / 1000.0  # DatasetReader.cpp:130, millimetres to metres
INTRINSICS_FILE = "intrinsics_kinect1.txt"


def _open_png(path):
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - Pillow is part of the image
        raise RuntimeError("reading dataset PNGs needs Pillow (PIL)") from e
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return Image.open(path)


def read_depth_png(path) -> np.ndarray:
    """readDepthImage (DatasetReader.cpp:122-134): 16-bit depth PNG -> float64 metres."""
    img = _open_png(path)
    a = np.array(img)
    if a.dtype not in (np.uint16, np.int32, np.uint8) or a.ndim != 2:
        raise ValueError(f"{path}: expected a single-channel 16-bit depth image, got {a.dtype} {a.shape}")
    return a.astype(np.float64) * DEPTH_SCALE


def read_gray_png(path) -> np.ndarray:
    return np.array(_open_png(path).convert("L"))


def read_color_png(path) -> np.ndarray:
    """Colour image as cv::imread returns it: 8-bit, channels in BGR order."""
    return np.array(_open_png(path).convert("RGB"))[:, :, ::-1].copy()


def load_matrix_from_file(path, m: int) -> list:
    """DatasetReader::LoadMatrixFromFile (DatasetReader.cpp:37-49): the first m numbers."""
    with open(path) as f:
        vals = f.read().split()
    if len(vals) < m:
        raise ValueError(f"{path}: expected {m} numbers, found {len(vals)}")
    return [float(v) for v in vals[:m]]


class DatasetReader:
    """DatasetReader (DatasetReader.h / .cpp) over a local copy of a Slavcheva sequence."""

    def __init__(self, root, dataset: str = "Snoopy"):
        if dataset not in DATASETS:
            raise ValueError(f"unknown dataset {dataset!r} (known: {sorted(DATASETS)})")
        self.root = str(root)
        self.dataset = dataset
        self.image_dir = os.path.join(self.root, dataset)
        self.K = np.array(load_matrix_from_file(os.path.join(self.root, INTRINSICS_FILE), 9)).reshape(3, 3)
        self.width, self.height = WIDTH, HEIGHT
        d = DATASETS[dataset]
        self.min_depth, self.max_depth = d["min_depth"], d["max_depth"]
        self.num_frames = d["frames"]

    def getNumImageFiles(self) -> int:
        return self.num_frames

    def getDepthIntrinsicMatrix(self) -> np.ndarray:
        return self.K

    def getDepthWidth(self) -> int:
        return self.width

    def getDepthHeight(self) -> int:
        return self.height

    def getMinimumDepthThreshold(self) -> float:
        return self.min_depth

    def getMaximumDepthThreshold(self) -> float:
        return self.max_depth

    def _file(self, kind: str, frame: int) -> str:
        return os.path.join(self.image_dir, f"{kind}_{frame:06d}.png")

    def getImages(self, frame: int):
        """getImages (DatasetReader.cpp:51-67): [colour (BGR uint8), masked depth (float64 m)]."""
        color = read_color_png(self._file("color", frame)) if os.path.exists(self._file("color", frame)) else None
        depth = read_depth_png(self._file("depth", frame))
        mask = read_gray_png(self._file("omask", frame))
        if mask.shape != depth.shape:
            raise ValueError(f"frame {frame}: mask {mask.shape} and depth {depth.shape} differ")
        return [color, np.where(mask != 0, depth, 0.0)]

    def depth_frame(self, frame: int) -> np.ndarray:
        """The masked depth as the float32 metres kf_integrate_depth takes."""
        return self.getImages(frame)[1].astype(np.float32)


def compute_bounds(K, w: int, h: int, min_depth: float, max_depth: float):
    """KillingFusion::computeBounds (KillingFusion.cpp:478-519): back-project the image corners
    (0, 0), (0, h-1), (w-1, h-1), (w-1, 0) at min and max depth through K^-1 (camera = world)
    and take the per-axis min and max."""
    K = np.asarray(K, dtype=np.float64).reshape(3, 3)
    corners = np.array([[0, 0, w - 1, w - 1, 0, 0, w - 1, w - 1],
                        [0, h - 1, h - 1, 0, 0, h - 1, h - 1, 0],
                        [1, 1, 1, 1, 1, 1, 1, 1]], dtype=np.float64)
    depth = np.array([min_depth] * 4 + [max_depth] * 4)
    rays = np.linalg.inv(K) @ corners
    pts = np.vstack([rays[:2] * depth, depth])
    return tuple(pts.min(axis=1)), tuple(pts.max(axis=1))


def pipeline_for(reader: DatasetReader, ctx, params: api.KillingParams, dtype: str = "f32"):
    """The reference's set-up for a dataset: one grid over the camera frustum between the
    dataset's depth limits at params.voxel_size (SDF::SDF with computeVoxelGridSize), identity
    pose.  Returns (per-frame driver, dims, min corner)."""
    mn, mx = compute_bounds(reader.K, reader.width, reader.height, reader.min_depth, reader.max_depth)
    dims = api.dims_from_bounds(mn, mx, params.voxel_size)
    pl = api.KillingFusion(ctx, dims, mn, params.voxel_size, params, reader.K, None, reader.min_depth,
                           reader.max_depth, reader.width, reader.height, dtype=dtype)
    return pl, dims, mn
