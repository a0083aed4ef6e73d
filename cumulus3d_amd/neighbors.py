"""Source-view selection and depth ranges from sparse point tracks (include/c3d_scene.h;
PairStats and depth_ranges, crates/dense/src/neighbors.rs:39-189): pair statistics and ranges on
the device, the per-view selection on the host side of the C ABI.

  count, q75 = pair_stats(backend, scene)
  sources = select_all(count, q75, max_views=10, min_angle_deg=1.0)   # densify.rs:239
  ranges = depth_ranges(backend, scene)                                # densify.rs:241
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import pm

EXPORTS = ["c3d_pair_stats", "c3d_select_views", "c3d_select_diverse", "c3d_depth_ranges"]


class c3d_sparse_scene(C.Structure):
    _fields_ = [
        ("num_views", C.c_uint32),
        ("pad0", C.c_uint32),
        ("rotation", C.POINTER(C.c_double)),
        ("translation", C.POINTER(C.c_double)),
        ("num_points", C.c_uint64),
        ("xyz", C.POINTER(C.c_double)),
        ("track_offset", C.POINTER(C.c_uint64)),
        ("track_views", C.POINTER(C.c_uint32)),
    ]


class SparseScene:
    """The ctypes image of a scene's cameras and sparse tracks, owning its arrays. `scene` is a
    scenes.DenseScene (views with r, t; points with xyz and views)."""

    def __init__(self, scene):
        n = len(scene.views)
        self.rot = np.ascontiguousarray(np.array([np.asarray(v.r, np.float64).reshape(9) for v in scene.views]), np.float64)
        self.trans = np.ascontiguousarray(np.array([np.asarray(v.t, np.float64).reshape(3) for v in scene.views]), np.float64)
        pts = scene.points
        self.xyz = np.ascontiguousarray(np.array([np.asarray(p.xyz, np.float64) for p in pts]).reshape(-1, 3), np.float64)
        self.off = np.zeros(len(pts) + 1, np.uint64)
        self.off[1:] = np.cumsum([len(p.views) for p in pts]) if pts else []
        self.tv = np.ascontiguousarray(np.array([v for p in pts for v in p.views], np.uint32).reshape(-1))
        self.c = c3d_sparse_scene()
        self.c.num_views = n
        self.c.rotation = self.rot.ctypes.data_as(C.POINTER(C.c_double))
        self.c.translation = self.trans.ctypes.data_as(C.POINTER(C.c_double))
        self.c.num_points = len(pts)
        self.c.xyz = self.xyz.ctypes.data_as(C.POINTER(C.c_double))
        self.c.track_offset = self.off.ctypes.data_as(C.POINTER(C.c_uint64))
        self.c.track_views = self.tv.ctypes.data_as(C.POINTER(C.c_uint32))
        self.n = n


def _declare(L):
    if getattr(L, "_c3d_scene_declared", False):
        return L
    u32p, dp = C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    sp = C.POINTER(c3d_sparse_scene)
    L.c3d_pair_stats.argtypes = [C.c_void_p, sp, u32p, dp]
    L.c3d_select_views.argtypes = [C.c_uint32, u32p, dp, C.c_uint32, C.c_uint32, C.c_double, u32p, u32p]
    L.c3d_select_diverse.argtypes = [sp, u32p, dp, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_double, u32p, u32p]
    L.c3d_depth_ranges.argtypes = [C.c_void_p, sp, dp, dp, C.POINTER(C.c_uint8)]
    for n in EXPORTS:
        getattr(L, n).restype = C.c_int
    L._c3d_scene_declared = True
    return L


def _sparse(scene) -> SparseScene:
    return scene if isinstance(scene, SparseScene) else SparseScene(scene)


def pair_stats(backend: pm.HipPatchMatch, scene) -> Tuple[np.ndarray, np.ndarray]:
    """PairStats::new on the device: (count, q75), n x n, symmetric."""
    L = _declare(backend.L)
    s = _sparse(scene)
    count = np.zeros((s.n, s.n), np.uint32)
    q75 = np.zeros((s.n, s.n), np.float64)
    pm._check(L, L.c3d_pair_stats(backend.h, C.byref(s.c), count.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  q75.ctypes.data_as(C.POINTER(C.c_double))))
    return count, q75


def select(count: np.ndarray, q75: np.ndarray, r: int, max_views: int, min_angle_rad: float) -> List[int]:
    """PairStats::select (neighbors.rs:83-103)."""
    L = _declare(pm.load())
    n = count.shape[0]
    count = np.ascontiguousarray(count, np.uint32)
    q75 = np.ascontiguousarray(q75, np.float64)
    out = (C.c_uint32 * max(max_views, 1))()
    m = C.c_uint32()
    pm._check(L, L.c3d_select_views(n, count.ctypes.data_as(C.POINTER(C.c_uint32)), q75.ctypes.data_as(C.POINTER(C.c_double)),
                                    r, max_views, min_angle_rad, out, C.byref(m)))
    return [int(out[i]) for i in range(m.value)]


def select_all(count: np.ndarray, q75: np.ndarray, max_views: int = 10, min_angle_deg: float = 1.0) -> List[List[int]]:
    """PairStats::select_all (neighbors.rs:154-157): NeighborParams defaults (params.rs:366-370)."""
    a = math.radians(min_angle_deg)
    return [select(count, q75, r, max_views, a) for r in range(count.shape[0])]


def select_diverse(scene, count: np.ndarray, q75: np.ndarray, r: int, max_views: int, min_angle_rad: float, bins: int = 8,
                   decay: float = 1.0) -> List[int]:
    """PairStats::select_diverse (neighbors.rs:110-151)."""
    L = _declare(pm.load())
    s = _sparse(scene)
    count = np.ascontiguousarray(count, np.uint32)
    q75 = np.ascontiguousarray(q75, np.float64)
    out = (C.c_uint32 * max(max_views, 1))()
    m = C.c_uint32()
    pm._check(L, L.c3d_select_diverse(C.byref(s.c), count.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      q75.ctypes.data_as(C.POINTER(C.c_double)), r, max_views, min_angle_rad, bins, decay, out,
                                      C.byref(m)))
    return [int(out[i]) for i in range(m.value)]


def depth_ranges(backend: pm.HipPatchMatch, scene) -> List[Optional[Tuple[float, float]]]:
    """depth_ranges (neighbors.rs:162-189) on the device: per view (0.75 lo, 1.25 hi) or None."""
    L = _declare(backend.L)
    s = _sparse(scene)
    lo = np.zeros(s.n, np.float64)
    hi = np.zeros(s.n, np.float64)
    has = np.zeros(s.n, np.uint8)
    pm._check(L, L.c3d_depth_ranges(backend.h, C.byref(s.c), lo.ctypes.data_as(C.POINTER(C.c_double)),
                                    hi.ctypes.data_as(C.POINTER(C.c_double)), has.ctypes.data_as(C.POINTER(C.c_uint8))))
    return [(float(lo[v]), float(hi[v])) if has[v] else None for v in range(s.n)]


def sources_and_ranges(backend: pm.HipPatchMatch, scene, max_views: int = 10, min_angle_deg: float = 1.0):
    """The neighbour step of compute_depth_maps (densify.rs:224-241, diversity off): per view its
    sources and its depth range."""
    s = _sparse(scene)
    count, q75 = pair_stats(backend, s)
    return select_all(count, q75, max_views, min_angle_deg), depth_ranges(backend, s)
