"""Agreement fusion of final depth maps into a point cloud on the MI355X (include/c3d_fusion.h;
fuse_consistency, crates/dense/src/fusion.rs:202-490), and the comparison of two clouds (median
point-to-point distance, completeness within 5 cm).

Fusion runs only on the device; the cloud comparison is an evaluation metric computed on the
host with a k-d tree (scipy).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import abi, pm, scenes

EXPORTS = ["c3d_fuse_consistency"]


class c3d_fusion_view(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("depth", C.POINTER(C.c_float)), ("normal", C.POINTER(C.c_float)),
                ("cost", C.POINTER(C.c_float)), ("gray", C.POINTER(C.c_uint8)), ("k", C.c_double * 4), ("r", C.c_double * 9),
                ("t", C.c_double * 3), ("baseline", C.c_double), ("image_id", C.c_uint32), ("num_overlap", C.c_uint32),
                ("overlap", C.POINTER(C.c_uint32))]


class c3d_fusion_params(C.Structure):
    _fields_ = [("min_consistent_views", C.c_uint32), ("consistency_num_images", C.c_uint32), ("mark_used", C.c_int32),
                ("mark_radius", C.c_uint32), ("inverse_variance", C.c_int32), ("pad0", C.c_uint32), ("sigma_px0", C.c_double),
                ("sigma_px_slope", C.c_double), ("max_reproj_error", C.c_double), ("max_depth_error", C.c_double),
                ("max_normal_error_deg", C.c_double)]


class c3d_point_cloud(C.Structure):
    _fields_ = [("capacity", C.c_uint64), ("count", C.c_uint64), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("colors", C.POINTER(C.c_uint8)), ("keys", C.POINTER(C.c_uint64))]


@dataclass
class FusionParams:
    """FusionParams (params.rs:236-290) with the fast profile's agreement tolerances
    (params.rs:456-458)."""

    min_consistent_views: int = 3
    consistency_num_images: int = 12
    mark_used: bool = True
    mark_radius: int = 0
    inverse_variance: bool = True
    sigma_px0: float = 0.25
    sigma_px_slope: float = 1.0
    max_reproj_error: float = 100.0
    max_depth_error: float = 0.005
    max_normal_error_deg: float = 23.0

    def c(self) -> c3d_fusion_params:
        return c3d_fusion_params(self.min_consistent_views, self.consistency_num_images, int(self.mark_used), self.mark_radius,
                                 int(self.inverse_variance), 0, self.sigma_px0, self.sigma_px_slope, self.max_reproj_error,
                                 self.max_depth_error, self.max_normal_error_deg)


@dataclass
class Cloud:
    positions: np.ndarray  # (n, 3) float32, world
    normals: np.ndarray
    colors: np.ndarray     # (n, 3) uint8
    keys: np.ndarray       # (n,) uint64


def baseline(scene: scenes.DenseScene, v: int, sources: Sequence[int]) -> float:
    """Median centre distance to the sources (densify.rs:260-276)."""
    c = scene.views[v].center()
    b = sorted(float(np.linalg.norm(c - scene.views[s].center())) for s in sources)
    if not b:
        return 0.0
    q = 0.5 * (len(b) - 1)
    lo = int(math.floor(q))
    hi = min(lo + 1, len(b) - 1)
    return b[lo] + (q - lo) * (b[hi] - b[lo])


def fusion_views(scene: scenes.DenseScene, depth: Sequence[np.ndarray], normal: Sequence[np.ndarray], cost: Sequence[np.ndarray],
                 overlap: Sequence[Sequence[int]], sources: Sequence[Sequence[int]], keep: list):
    """c3d_fusion_view array for every view of `scene` (depth 0 = invalid)."""
    arr = (c3d_fusion_view * len(scene.views))()
    for v, view in enumerate(scene.views):
        d = np.ascontiguousarray(depth[v], np.float32)
        n = np.ascontiguousarray(normal[v], np.float32)
        c = np.ascontiguousarray(cost[v], np.float32) if cost is not None else None
        g = np.ascontiguousarray(view.gray, np.uint8)
        ov = np.ascontiguousarray(np.asarray(overlap[v], np.uint32))
        keep += [d, n, c, g, ov]
        f = arr[v]
        f.height, f.width = d.shape
        f.depth = d.ctypes.data_as(C.POINTER(C.c_float))
        f.normal = n.ctypes.data_as(C.POINTER(C.c_float))
        f.cost = c.ctypes.data_as(C.POINTER(C.c_float)) if c is not None else None
        f.gray = g.ctypes.data_as(C.POINTER(C.c_uint8))
        f.k[:] = [float(q) for q in view.k]
        f.r[:] = [float(q) for q in np.asarray(view.r, float).reshape(9)]
        f.t[:] = [float(q) for q in view.t]
        f.baseline = baseline(scene, v, sources[v])
        f.image_id = v
        f.num_overlap = len(ov)
        f.overlap = ov.ctypes.data_as(C.POINTER(C.c_uint32)) if len(ov) else None
    return arr


def _cloud(cap):
    pos = np.zeros((cap, 3), np.float32)
    nrm = np.zeros((cap, 3), np.float32)
    col = np.zeros((cap, 3), np.uint8)
    key = np.zeros(cap, np.uint64)
    pc = c3d_point_cloud(cap, 0, pos.ctypes.data_as(C.POINTER(C.c_float)), nrm.ctypes.data_as(C.POINTER(C.c_float)),
                         col.ctypes.data_as(C.POINTER(C.c_uint8)), key.ctypes.data_as(C.POINTER(C.c_uint64)))
    return pc, (pos, nrm, col, key)


def fuse(backend: pm.HipPatchMatch, views_arr, params: FusionParams, capacity: int) -> Cloud:
    L = backend.L
    if not getattr(L, "_c3d_fusion_declared", False):
        L.c3d_fuse_consistency.argtypes = [C.c_void_p, C.POINTER(c3d_fusion_view), C.c_uint32, C.POINTER(c3d_fusion_params),
                                           C.POINTER(c3d_point_cloud)]
        L.c3d_fuse_consistency.restype = C.c_int
        L._c3d_fusion_declared = True
    pc, arrs = _cloud(capacity)
    p = params.c()
    pm._check(L, L.c3d_fuse_consistency(backend.h, views_arr, len(views_arr), C.byref(p), C.byref(pc)))
    m = int(pc.count)
    return Cloud(arrs[0][:m], arrs[1][:m], arrs[2][:m], arrs[3][:m])


def compare(a: Cloud, b: Cloud, radius: float = 0.05, max_queries: int = 2_000_000):
    """Median distance from each point of `a` to the nearest point of `b`, and the completeness of
    `b`: the fraction of its points with a point of `a` within `radius` (m). Above `max_queries`
    points the query side is a seeded uniform sample (the trees hold every point)."""
    from scipy.spatial import cKDTree

    if len(a.positions) == 0 or len(b.positions) == 0:
        return float("inf"), 0.0
    rng = np.random.default_rng(0)
    qa = a.positions if len(a.positions) <= max_queries else a.positions[rng.choice(len(a.positions), max_queries, replace=False)]
    qb = b.positions if len(b.positions) <= max_queries else b.positions[rng.choice(len(b.positions), max_queries, replace=False)]
    da, _ = cKDTree(b.positions.astype(np.float64)).query(qa.astype(np.float64), k=1, workers=16)
    db, _ = cKDTree(a.positions.astype(np.float64)).query(qb.astype(np.float64), k=1, workers=16, distance_upper_bound=radius)
    return float(np.median(da)), float(np.mean(np.isfinite(db)))
