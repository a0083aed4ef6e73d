"""Python binding of the learned initialiser (include/c3d_init.h): a bf16 U-Net on the gfx950
matrix cores that predicts depth, normal and confidence from plane-sweep features, trained with
hand-written forward, backward and Adam kernels on plain-PatchStereo labels.

There is no CPU path: every call runs the HIP kernels of libc3d_pm.so and raises without them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import abi, pm

EXPORTS = ["c3d_init_create", "c3d_init_destroy", "c3d_init_num_params", "c3d_init_get_params", "c3d_init_set_params",
           "c3d_init_features", "c3d_init_set_features", "c3d_init_get_features", "c3d_init_set_labels", "c3d_init_train_step", "c3d_init_predict",
           "c3d_pm_densify_learned"]


class c3d_sweep_source(C.Structure):
    _fields_ = [("gray", C.POINTER(C.c_uint8)), ("width", C.c_uint32), ("height", C.c_uint32), ("k", C.c_float * 4),
                ("r", C.c_float * 9), ("t", C.c_float * 3)]


class c3d_sweep_view(C.Structure):
    _fields_ = [("gray", C.POINTER(C.c_uint8)), ("width", C.c_uint32), ("height", C.c_uint32), ("k", C.c_float * 4),
                ("num_sources", C.c_uint32), ("sources", c3d_sweep_source * 2), ("depth_min", C.c_float),
                ("depth_max", C.c_float)]


class c3d_init_train_params(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("w_normal", C.c_float),
                ("w_conf", C.c_float), ("conf_tol", C.c_float), ("pad0", C.c_uint32)]


class c3d_learned_schedule(C.Structure):
    _fields_ = [("refine_iters", C.c_uint32), ("geometric_rounds", C.c_uint32), ("geometric_iters", C.c_uint32),
                ("first_run_id", C.c_uint32), ("conf_min", C.c_float), ("pad0", C.c_uint32),
                ("post", abi.c3d_post_params), ("post_k", C.POINTER(C.c_double))]


@dataclass
class LearnedSchedule:
    """PatchStereo after the learned initialisation: refinement iterations, geometric rounds,
    confidence below which the reference's random initialisation is used."""

    refine_iters: int = 2
    geometric_rounds: int = 1
    geometric_iters: int = 2
    first_run_id: int = 1
    conf_min: float = 0.5
    post: Optional[abi.PostParams] = None  # read-out post-processing, as in abi.Schedule

    def c(self, view_k=None) -> c3d_learned_schedule:
        sch = c3d_learned_schedule(self.refine_iters, self.geometric_rounds, self.geometric_iters, self.first_run_id, self.conf_min, 0)
        if self.post is not None:
            sch.post = self.post.c()
        if view_k is not None:
            sch._k = abi.post_k_array(view_k)
            sch.post_k = C.cast(sch._k, C.POINTER(C.c_double))
        return sch


@dataclass
class TrainParams:
    lr: float = 1e-3
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    w_normal: float = 0.5
    w_conf: float = 0.1
    conf_tol: float = 0.01

    def c(self) -> c3d_init_train_params:
        return c3d_init_train_params(self.lr, self.beta1, self.beta2, self.eps, self.w_normal, self.w_conf, self.conf_tol, 0)


def _declare(L):
    if getattr(L, "_c3d_init_declared", False):
        return L
    fp = C.POINTER(C.c_float)
    L.c3d_init_create.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]
    L.c3d_init_destroy.argtypes = [C.c_void_p]
    L.c3d_init_destroy.restype = None
    L.c3d_init_num_params.argtypes = [C.c_void_p]
    L.c3d_init_num_params.restype = C.c_uint64
    L.c3d_init_get_params.argtypes = [C.c_void_p, fp]
    L.c3d_init_set_params.argtypes = [C.c_void_p, fp]
    L.c3d_init_features.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(c3d_sweep_view)]
    L.c3d_init_set_features.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint16)]
    L.c3d_init_get_features.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint16)]
    L.c3d_init_set_labels.argtypes = [C.c_void_p, C.c_uint32, fp, fp, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
    L.c3d_init_train_step.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(c3d_init_train_params), C.c_int, fp, fp]
    L.c3d_init_predict.argtypes = [C.c_void_p, C.c_uint32, fp, fp, fp, fp]
    fpp = C.POINTER(C.POINTER(C.c_float))
    L.c3d_pm_densify_learned.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(c3d_learned_schedule),
                                         C.POINTER(abi.c3d_view_state), C.POINTER(C.POINTER(C.c_uint8)), fpp, fpp, fpp,
                                         C.POINTER(abi.c3d_timing)]
    L.c3d_pm_densify_learned.restype = C.c_int
    for n in ("c3d_init_create", "c3d_init_get_params", "c3d_init_set_params", "c3d_init_features", "c3d_init_set_features", "c3d_init_get_features",
              "c3d_init_set_labels", "c3d_init_train_step", "c3d_init_predict"):
        getattr(L, n).restype = C.c_int
    L._c3d_init_declared = True
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def sweep_view(gray: np.ndarray, k: Sequence[float], sources: Sequence[tuple], dmin: float, dmax: float, keep: list):
    """c3d_sweep_view of a reference image and up to 2 sources given as (gray, k, r[9], t[3])."""
    v = c3d_sweep_view()
    g = np.ascontiguousarray(gray, np.uint8)
    keep.append(g)
    v.gray = g.ctypes.data_as(C.POINTER(C.c_uint8))
    v.height, v.width = g.shape
    v.k[:] = [float(q) for q in k]
    v.num_sources = len(sources)
    for j, (sg, sk, r, t) in enumerate(sources):
        sg = np.ascontiguousarray(sg, np.uint8)
        keep.append(sg)
        s = v.sources[j]
        s.gray = sg.ctypes.data_as(C.POINTER(C.c_uint8))
        s.height, s.width = sg.shape
        s.k[:] = [float(q) for q in sk]
        s.r[:] = [float(q) for q in np.asarray(r).reshape(9)]
        s.t[:] = [float(q) for q in t]
    v.depth_min = dmin
    v.depth_max = dmax
    return v


class InitNet:
    """The U-Net on the device of `backend` for batches of up to max_batch images of height x width."""

    def __init__(self, backend: pm.HipPatchMatch, max_batch: int, height: int, width: int, seed: int = 1):
        self.L = _declare(backend.L)
        self.be = backend
        self.B, self.H, self.W = max_batch, height, width
        h = C.c_void_p()
        pm._check(self.L, self.L.c3d_init_create(backend.h, max_batch, height, width, seed, C.byref(h)))
        self.h = h
        self.n = int(self.L.c3d_init_num_params(self.h))
        backend._adopt(self)

    def close(self):
        if self.h:
            self.L.c3d_init_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        pm._check(self.L, self.L.c3d_init_get_params(self.h, _f(out)))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        pm._check(self.L, self.L.c3d_init_set_params(self.h, _f(p)))

    def features(self, slot0: int, views: Sequence[c3d_sweep_view]):
        arr = (c3d_sweep_view * len(views))(*views)
        pm._check(self.L, self.L.c3d_init_features(self.h, slot0, len(views), arr))

    def set_features(self, feat_bits: np.ndarray):
        f = np.ascontiguousarray(feat_bits, np.uint16)
        pm._check(self.L, self.L.c3d_init_set_features(self.h, f.shape[0], f.ctypes.data_as(C.POINTER(C.c_uint16))))

    def get_features(self, slot: int) -> np.ndarray:
        out = np.zeros((self.H, self.W, 64), np.uint16)
        pm._check(self.L, self.L.c3d_init_get_features(self.h, slot, out.ctypes.data_as(C.POINTER(C.c_uint16))))
        return out

    def set_labels(self, slot: int, depth: np.ndarray, normal: np.ndarray, valid_w: int, valid_h: int, dmin: float, dmax: float):
        d = np.ascontiguousarray(depth, np.float32)
        nrm = np.ascontiguousarray(normal, np.float32)
        assert d.shape == (self.H, self.W) and nrm.shape == (self.H, self.W, 3)
        pm._check(self.L, self.L.c3d_init_set_labels(self.h, slot, _f(d), _f(nrm), valid_w, valid_h, dmin, dmax))

    def train_step(self, batch: int, params: TrainParams, step: bool = True, want_grads: bool = False):
        losses = np.zeros(4, np.float32)
        grads = np.zeros(self.n, np.float32) if want_grads else None
        tp = params.c()
        pm._check(self.L, self.L.c3d_init_train_step(self.h, batch, C.byref(tp), int(step), _f(losses), _f(grads)))
        return losses, grads

    def predict(self, batch: int, want_head: bool = False):
        depth = np.zeros((batch, self.H, self.W), np.float32)
        normal = np.zeros((batch, self.H, self.W, 3), np.float32)
        conf = np.zeros((batch, self.H, self.W), np.float32)
        head = np.zeros((batch, self.H, self.W, 5), np.float32) if want_head else None
        pm._check(self.L, self.L.c3d_init_predict(self.h, batch, _f(depth), _f(normal), _f(conf), _f(head)))
        return depth, normal, conf, head


def output_buffers(inp, views: Sequence[int]):
    """(states, counts, raw depth) arrays for the top-scale results of `views`, for densify_learned's
    `out` (a caller that keeps them between jobs)."""
    top = inp.num_levels - 1
    outs, counts, raws = [], [], []
    for v in views:
        h, w = inp.level_shape(v, top)
        outs.append(abi.ViewState(w, h))
        counts.append(np.zeros((h, w), np.uint8))
        raws.append(np.zeros((h, w), np.float32))
    return outs, counts, raws


def densify_learned(session: pm.Session, net: InitNet, views: Sequence[int], schedule: LearnedSchedule, download: bool = True,
                    want_init: bool = False, out=None):
    """c3d_pm_densify_learned: returns (states, counts, raw depth, timing, init depth, init normal)
    at the top scale; the arrays are None without download / want_init. `out`: arrays from
    output_buffers to download into instead of new ones."""
    L = net.L
    inp = session.inp
    top = inp.num_levels - 1
    sch = schedule.c([inp.view_k[v] for v in views] if schedule.post is not None else None)
    tm = abi.c3d_timing()
    outs = counts = raws = idep = inor = None
    if download:
        outs, counts, raws = out if out is not None else output_buffers(inp, views)
        if len(outs) != len(views) or len(counts) != len(views) or len(raws) != len(views):
            raise ValueError("out holds arrays for a different number of views")
    if want_init:
        idep, inor = [], []
        for v in views:
            h, w = inp.level_shape(v, top)
            idep.append(np.zeros((h, w), np.float32))
            inor.append(np.zeros((h, w, 3), np.float32))
    rc = L.c3d_pm_densify_learned(session.h, net.h, abi.views_array(views), len(views), C.byref(sch),
                                  abi.states_array(outs) if outs else None,
                                  abi.ptr_array(counts, C.c_uint8) if counts else None,
                                  abi.ptr_array(raws, C.c_float) if raws else None,
                                  abi.ptr_array(idep, C.c_float) if idep else None,
                                  abi.ptr_array(inor, C.c_float) if inor else None, C.byref(tm))
    pm._check(L, rc)
    return outs, counts, raws, tm, idep, inor
