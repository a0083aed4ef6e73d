"""Python binding of the gfx950 PatchStereo engine (include/c3d_pm.h) with the reference's interface.

``HipPatchMatch`` stands in for ``CudaPatchMatch`` behind ``PatchMatchBackend``
(crates/cuda/src/patchmatch.rs:285-298, crates/dense/src/kernel.rs:239-245) and ``Session`` for
``PatchMatchSession`` (kernel.rs:191-237): ``run``, ``evaluate``, ``filter``, ``upsample``,
``median_filter`` with the same argument meaning. Errors raise :class:`PmError` carrying the
status the Rust shim maps to ``Error::InvalidArgument`` / ``Error::Invariant`` (patchmatch.rs:234-236).

There is no CPU fallback: without the built library or without a gfx950 device every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
import weakref
from typing import List, Optional, Sequence

import numpy as np

from . import abi

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_libs = {}
_lock = threading.Lock()

EXPORTS = [
    "c3d_pm_create", "c3d_pm_name", "c3d_pm_destroy", "c3d_pm_is_available", "c3d_pm_begin", "c3d_pm_run",
    "c3d_pm_evaluate", "c3d_pm_filter", "c3d_pm_upsample", "c3d_pm_median", "c3d_pm_end", "c3d_pm_last_error",
    "c3d_pm_densify", "c3d_pm_pair_costs", "c3d_pm_num_levels", "c3d_pm_build_pyramid", "c3d_pm_autotune",
    "c3d_remove_speckles", "c3d_fill_holes", "c3d_fill_holes_noalloc", "c3d_pm_sample_reciprocal",
]


class PmError(RuntimeError):
    """A failed engine call: ``status`` is the c3d_status code, the message the engine's text."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message


def library_path() -> str:
    """The engine library: $C3D_PM_LIBRARY when set (another build of the same ABI), else the
    in-tree build."""
    return os.environ.get("C3D_PM_LIBRARY") or os.path.join(_LIB_DIR, "libc3d_pm.so")


def load() -> C.CDLL:
    """Loads the engine library (built by cumulus3d_amd/build.py); raises if it is missing."""
    with _lock:
        if "engine" in _libs:
            return _libs["engine"]
        path = library_path()
        if not os.path.exists(path):
            raise PmError(abi.C3D_NOT_AVAILABLE, f"{path} is missing: build the extension first (__graft_entry__.build())")
        L = C.CDLL(path)
        u32p = C.POINTER(C.c_uint32)
        vsp = C.POINTER(abi.c3d_view_state)
        snp = C.POINTER(abi.c3d_depth_snapshot)
        L.c3d_pm_create.argtypes = [C.c_int, C.POINTER(abi.c3d_pm_options), C.POINTER(C.c_void_p)]
        L.c3d_pm_name.argtypes = [C.c_void_p]
        L.c3d_pm_name.restype = C.c_char_p
        L.c3d_pm_destroy.argtypes = [C.c_void_p]
        L.c3d_pm_destroy.restype = C.c_int
        L.c3d_pm_is_available.argtypes = []
        L.c3d_pm_begin.argtypes = [C.c_void_p, C.POINTER(abi.c3d_kernel_input), C.POINTER(C.c_void_p)]
        L.c3d_pm_run.argtypes = [C.c_void_p, u32p, C.c_uint32, vsp, C.POINTER(abi.c3d_run_params), snp]
        L.c3d_pm_evaluate.argtypes = [C.c_void_p, C.c_uint32, u32p, C.c_uint32, vsp, C.c_int, snp,
                                      C.POINTER(C.POINTER(C.c_float))]
        L.c3d_pm_filter.argtypes = [C.c_void_p, u32p, C.c_uint32, vsp, snp, C.POINTER(C.POINTER(C.c_uint8))]
        L.c3d_pm_upsample.argtypes = [C.c_void_p, C.c_uint32, u32p, C.c_uint32, vsp, C.c_float, C.c_float, vsp]
        L.c3d_pm_median.argtypes = [C.c_void_p, C.c_uint32, u32p, C.c_uint32, vsp, vsp]
        L.c3d_pm_end.argtypes = [C.c_void_p]
        L.c3d_pm_end.restype = None
        L.c3d_pm_last_error.restype = C.c_char_p
        L.c3d_pm_densify.argtypes = [C.c_void_p, u32p, C.c_uint32, C.POINTER(abi.c3d_schedule), vsp,
                                     C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.POINTER(C.c_float)),
                                     C.POINTER(abi.c3d_timing)]
        L.c3d_pm_pair_costs.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.c3d_pm_autotune.argtypes = [C.c_void_p, u32p, C.c_uint32, C.c_uint32, u32p, u32p, C.POINTER(C.c_double)]
        L.c3d_pm_autotune.restype = C.c_int
        L.c3d_pm_num_levels.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.c3d_pm_num_levels.restype = C.c_uint32
        L.c3d_pm_build_pyramid.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32,
                                           C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.POINTER(C.c_uint8)),
                                           C.POINTER(C.c_float)]
        fp = C.POINTER(C.c_float)
        L.c3d_remove_speckles.argtypes = [C.c_void_p, fp, fp, fp, C.c_uint32, C.c_uint32, C.POINTER(abi.c3d_post_params),
                                          C.POINTER(C.c_uint64)]
        L.c3d_fill_holes.argtypes = [C.c_void_p, fp, fp, fp, fp, C.POINTER(C.c_uint8), C.POINTER(C.c_double), C.c_uint32,
                                     C.c_uint32, C.POINTER(abi.c3d_post_params), C.POINTER(C.c_uint64)]
        L.c3d_fill_holes_noalloc.argtypes = L.c3d_fill_holes.argtypes
        L.c3d_pm_sample_reciprocal.argtypes = [C.c_void_p, fp, fp, C.c_uint64]
        for n in ("c3d_pm_create", "c3d_pm_begin", "c3d_pm_run", "c3d_pm_evaluate", "c3d_pm_filter", "c3d_pm_upsample",
                  "c3d_pm_median", "c3d_pm_densify", "c3d_pm_pair_costs", "c3d_pm_is_available", "c3d_pm_build_pyramid",
                  "c3d_remove_speckles", "c3d_fill_holes", "c3d_fill_holes_noalloc", "c3d_fill_holes_noalloc"):
            getattr(L, n).restype = C.c_int
        _libs["engine"] = L
        return L


def _check(L, rc: int):
    if rc != 0:
        raise PmError(rc, L.c3d_pm_last_error().decode(errors="replace"))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _post_shape(depth: np.ndarray, normal: np.ndarray, cost: np.ndarray):
    """(h, w) of a post-processing call; the arrays are written in place, so they must already be
    contiguous float32."""
    if depth.ndim != 2:
        raise ValueError("depth must be (h, w)")
    h, w = depth.shape
    for a, shape in ((depth, (h, w)), (normal, (h, w, 3)), (cost, (h, w))):
        if a.shape != shape or a.dtype != np.float32 or not a.flags.c_contiguous:
            raise ValueError(f"expected a contiguous float32 array of shape {shape}")
    return h, w


def num_levels(min_side: int, max_levels: int, min_size: int) -> int:
    """num_levels (crates/dense/src/upsample.rs:36-44)."""
    return int(load().c3d_pm_num_levels(min_side, max_levels, min_size))


def is_available() -> bool:
    """is_available (crates/cuda/src/device.rs:80-89): a gfx950 device can be opened."""
    return bool(load().c3d_pm_is_available())


class HipPatchMatch:
    """PatchMatchBackend over device `ordinal` (CudaPatchMatch::new / try_default, patchmatch.rs:239-246)."""

    def __init__(self, ordinal: int = 0, options: Optional[abi.c3d_pm_options] = None):
        self.L = load()
        h = C.c_void_p()
        _check(self.L, self.L.c3d_pm_create(ordinal, C.byref(options) if options is not None else None, C.byref(h)))
        self.h = h
        # sessions and networks on this device: released before the device handle, whatever
        # order the interpreter finalises objects in
        self._dependents = weakref.WeakSet()

    def _adopt(self, obj):
        self._dependents.add(obj)

    def name(self) -> str:
        return self.L.c3d_pm_name(self.h).decode()

    def begin(self, inp: abi.KernelInput) -> "Session":
        return Session(self, inp)

    def build_pyramid(self, gray: np.ndarray, k: Sequence[float], levels: int) -> List[abi.LevelData]:
        """Gray pyramid on the device (build_pyramid, crates/dense/src/upsample.rs:78-102): the
        levels coarsest first, each with its intrinsics."""
        gray = np.ascontiguousarray(gray, np.uint8)
        h, w = gray.shape
        outs = [np.zeros((h >> (levels - 1 - l), w >> (levels - 1 - l)), np.uint8) for l in range(levels)]
        kk = (C.c_double * 4)(*[float(q) for q in k])
        ko = (C.c_float * (4 * levels))()
        _check(self.L, self.L.c3d_pm_build_pyramid(self.h, gray.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, kk, levels,
                                                   abi.ptr_array(outs, C.c_uint8), ko))
        return [abi.LevelData(outs[l], [float(ko[4 * l + j]) for j in range(4)]) for l in range(levels)]

    def remove_speckles(self, depth: np.ndarray, normal: np.ndarray, cost: np.ndarray, params: "abi.PostParams") -> int:
        """Small-component removal on the device (c3d_remove_speckles; remove_speckles,
        crates/dense/src/postproc.rs:77-133), depth (h, w) and normal (h, w, 3) float32 in place.
        Returns the number of erased pixels."""
        h, w = _post_shape(depth, normal, cost)
        n = C.c_uint64()
        p = params.c()
        _check(self.L, self.L.c3d_remove_speckles(self.h, _fp(depth), _fp(normal), _fp(cost), w, h, C.byref(p), C.byref(n)))
        return int(n.value)

    def fill_holes(self, depth: np.ndarray, normal: np.ndarray, cost: np.ndarray, raw_cost: np.ndarray, gray: np.ndarray,
                   k: Sequence[float], params: "abi.PostParams") -> int:
        """Boundary-aware gap filling on the device (c3d_fill_holes; fill_holes,
        postproc.rs:135-204): depth, normal and cost in place. Returns the number of filled pixels."""
        h, w = _post_shape(depth, normal, cost)
        if raw_cost.shape != (h, w) or raw_cost.dtype != np.float32 or not raw_cost.flags.c_contiguous:
            raise ValueError("raw_cost must be a contiguous float32 (h, w) array")
        gray = np.ascontiguousarray(gray, np.uint8)
        if gray.shape != (h, w):
            raise ValueError("gray must be (h, w)")
        n = C.c_uint64()
        p = params.c()
        kk = (C.c_double * 4)(*[float(q) for q in k])
        _check(self.L, self.L.c3d_fill_holes(self.h, _fp(depth), _fp(normal), _fp(cost), _fp(raw_cost),
                                             gray.ctypes.data_as(C.POINTER(C.c_uint8)), kk, w, h, C.byref(p), C.byref(n)))
        return int(n.value)

    def sample_reciprocal(self, z: np.ndarray) -> np.ndarray:
        """The kernels' per-sample reciprocal 1/z on the device (c3d_pm_sample_reciprocal)."""
        z = np.ascontiguousarray(z, np.float32)
        r = np.zeros_like(z)
        _check(self.L, self.L.c3d_pm_sample_reciprocal(self.h, _fp(z), _fp(r), z.size))
        return r

    def close(self):
        if self.h:
            for d in list(self._dependents):
                d.close()
            self.L.c3d_pm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Session:
    """PatchMatchSession (kernel.rs:191-237) over the engine; images are uploaded at creation."""

    def __init__(self, backend: HipPatchMatch, inp: abi.KernelInput):
        self.be = backend
        self.L = backend.L
        self.inp = inp
        h = C.c_void_p()
        _check(self.L, self.L.c3d_pm_begin(backend.h, C.byref(inp.c), C.byref(h)))
        self.h = h
        backend._adopt(self)

    def close(self):
        if self.h:
            self.L.c3d_pm_end(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, views: Sequence[int], states: Sequence[abi.ViewState], params: abi.c3d_run_params,
            snapshot: Optional[abi.DepthSnapshot] = None) -> None:
        arr = abi.states_array(states)
        _check(self.L, self.L.c3d_pm_run(self.h, abi.views_array(views), len(views), arr, C.byref(params),
                                         C.byref(snapshot.c) if snapshot else None))

    def evaluate(self, level: int, views: Sequence[int], states: Sequence[abi.ViewState], geometric: bool = False,
                 snapshot: Optional[abi.DepthSnapshot] = None) -> List[np.ndarray]:
        outs = [np.zeros((s.height, s.width), np.float32) for s in states]
        _check(self.L, self.L.c3d_pm_evaluate(self.h, level, abi.views_array(views), len(views), abi.states_array(states),
                                              int(geometric), C.byref(snapshot.c) if snapshot else None,
                                              abi.ptr_array(outs, C.c_float)))
        return outs

    def filter(self, views: Sequence[int], states: Sequence[abi.ViewState], snapshot: abi.DepthSnapshot) -> List[np.ndarray]:
        outs = [np.zeros((s.height, s.width), np.uint8) for s in states]
        _check(self.L, self.L.c3d_pm_filter(self.h, abi.views_array(views), len(views), abi.states_array(states),
                                            C.byref(snapshot.c), abi.ptr_array(outs, C.c_uint8)))
        return outs

    def upsample(self, level: int, views: Sequence[int], low: Sequence[abi.ViewState], sigma_s: float,
                 sigma_c: float) -> List[abi.ViewState]:
        outs = []
        for v in views:
            h, w = self.inp.level_shape(v, level)
            outs.append(abi.ViewState(w, h))
        _check(self.L, self.L.c3d_pm_upsample(self.h, level, abi.views_array(views), len(views), abi.states_array(low),
                                              sigma_s, sigma_c, abi.states_array(outs)))
        return outs

    def median_filter(self, level: int, views: Sequence[int], states: Sequence[abi.ViewState]) -> List[abi.ViewState]:
        outs = [abi.ViewState(s.width, s.height) for s in states]
        _check(self.L, self.L.c3d_pm_median(self.h, level, abi.views_array(views), len(views), abi.states_array(states),
                                            abi.states_array(outs)))
        for o, s in zip(outs, states):
            o.cost[...] = s.cost
        return outs

    def densify(self, views: Sequence[int], schedule: abi.Schedule, download: bool = True):
        """Device-resident compute_depth_maps schedule over every scale (densify.rs:350-454). Returns
        (top-scale states with thresholded depth, counts, raw depth, timing), or (None, None, None,
        timing) without download."""
        sch = schedule.c([self.inp.view_k[v] for v in views] if schedule.post is not None else None)
        top = self.inp.num_levels - 1
        tm = abi.c3d_timing()
        if not download:
            _check(self.L, self.L.c3d_pm_densify(self.h, abi.views_array(views), len(views), C.byref(sch), None, None, None,
                                                 C.byref(tm)))
            return None, None, None, tm
        outs, counts, raws = [], [], []
        for v in views:
            h, w = self.inp.level_shape(v, top)
            outs.append(abi.ViewState(w, h))
            counts.append(np.zeros((h, w), np.uint8))
            raws.append(np.zeros((h, w), np.float32))
        _check(self.L, self.L.c3d_pm_densify(self.h, abi.views_array(views), len(views), C.byref(sch), abi.states_array(outs),
                                             abi.ptr_array(counts, C.c_uint8), abi.ptr_array(raws, C.c_float), C.byref(tm)))
        return outs, counts, raws, tm

    def autotune(self, views: Sequence[int], level: int = 0):
        """c3d_pm_autotune: the half-step shape kept for this session, (rows, lanes per pixel), and
        the timings of the shapes of abi.TILE_SHAPES in that order."""
        rows, tpp = C.c_uint32(), C.c_uint32()
        ms = (C.c_double * len(abi.TILE_SHAPES))()
        _check(self.L, self.L.c3d_pm_autotune(self.h, abi.views_array(views), len(views), level, C.byref(rows), C.byref(tpp), ms))
        return (int(rows.value), int(tpp.value)), tuple(ms)

    def pair_costs(self, level: int, view: int, slot: int, xy: np.ndarray, planes: np.ndarray) -> np.ndarray:
        """Per-hypothesis photometric cost of explicit planes (c3d_pm_pair_costs)."""
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
        planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
        out = np.zeros(len(xy), np.float32)
        _check(self.L, self.L.c3d_pm_pair_costs(self.h, level, view, slot, len(xy), xy.ctypes.data_as(C.POINTER(C.c_int32)),
                                                planes.ctypes.data_as(C.POINTER(C.c_float)),
                                                out.ctypes.data_as(C.POINTER(C.c_float))))
        return out
