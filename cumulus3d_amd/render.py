"""The terrain renderer on the device (libc3d_render.so, csrc/terrain.hip): the scene model of the
host renderer (csrc/terrain_core.h) with one GPU thread per pixel, for workloads too large to
render on the host in reasonable time (the 108-view 1080p zone). Test fixtures use the host
renderer, whose output the golden files are tied to; the two agree to the last bits except where
the host compiler fused a multiply-add differently.
"""
from __future__ import annotations

import ctypes as C
import os

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libc3d_render.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        lib.c3d_render_terrain_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                  C.POINTER(C.c_float)]
        lib.c3d_render_terrain_device.restype = C.c_int
        _LIB = lib
    return _LIB


def render_terrain_device(params, cams, n, img, dep, nor):
    """Renders `n` cameras (scenes._TerrainCamera array) with scenes._TerrainParams `params` into the
    float32 arrays img (n, h, w), dep (n, h, w), nor (n, h, w, 3)."""
    fp = C.POINTER(C.c_float)
    rc = _lib().c3d_render_terrain_device(C.byref(params), cams, n, img.ctypes.data_as(fp), dep.ctypes.data_as(fp),
                                          nor.ctypes.data_as(fp))
    if rc:
        raise RuntimeError(f"c3d_render_terrain_device failed with HIP error {rc}")
