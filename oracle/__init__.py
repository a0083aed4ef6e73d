"""CPU oracle for the PatchStereo hot path — TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker
(or the timed CPU baseline). The product (cumulus3d_amd) never imports this package.

``Oracle`` wraps oracle/build/libpm_oracle.so (built from pm_oracle.c by oracle/Makefile), a C
restatement of the reference's CUDA PatchStereo kernels and host rules; see pm_oracle.c for the
function-by-function citations and what pins each part.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

from cumulus3d_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libpm_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compiles the oracle with its Makefile (gcc)."""
    cmd = ["make", "-C", _HERE]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load_library(LIB_PATH)
    return _lib


def load_library(path: str):
    """Loads an oracle build at `path` and declares its signatures."""
    if True:
        L = C.CDLL(path)
        u32p = C.POINTER(C.c_uint32)
        vsp = C.POINTER(abi.c3d_view_state)
        L.pmo_hash_bytes.argtypes = [C.c_char_p, C.c_size_t]
        L.pmo_hash_bytes.restype = C.c_uint64
        L.pmo_mix64.argtypes = [C.c_uint64]
        L.pmo_mix64.restype = C.c_uint64
        L.pmo_rng_state.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        L.pmo_rng_state.restype = C.c_uint64
        L.pmo_rng_next.argtypes = [C.POINTER(C.c_uint64)]
        L.pmo_rng_next.restype = C.c_float
        for n in ("pmo_erf", "pmo_emission_norm"):
            getattr(L, n).argtypes = [C.c_double]
            getattr(L, n).restype = C.c_double
        for n in ("pmo_emission", "pmo_visibility_probability", "pmo_triangulation_prior", "pmo_incident_prior"):
            getattr(L, n).argtypes = [C.c_double, C.c_double]
            getattr(L, n).restype = C.c_double
        L.pmo_resolution_prior.argtypes = [C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double]
        L.pmo_resolution_prior.restype = C.c_double
        dp = C.POINTER(C.c_double)
        L.pmo_plane_transfer.argtypes = [C.c_double, dp, dp, dp, dp]
        L.pmo_plane_transfer.restype = C.c_int
        L.pmo_quantile_sorted.argtypes = [dp, C.c_size_t, C.c_double]
        L.pmo_quantile_sorted.restype = C.c_double
        L.pmo_bilinear_clamped.argtypes = [C.POINTER(C.c_float), C.c_size_t, C.c_size_t, C.c_float, C.c_float]
        L.pmo_bilinear_clamped.restype = C.c_float
        L.pmo_set_rcp_mode.argtypes = [C.c_int]
        L.pmo_rcp_sample.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t]
        L.pmo_expf.argtypes = [C.c_float]
        L.pmo_expf.restype = C.c_float
        L.pmo_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        fp = C.POINTER(C.c_float)
        L.pmo_level_pair.argtypes = [fp, fp, fp, fp, fp, fp]
        L.pmo_num_levels.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.pmo_num_levels.restype = C.c_uint32
        L.pmo_build_pyramid.argtypes = [C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32, dp, C.c_uint32,
                                        C.POINTER(C.POINTER(C.c_uint8)), fp]
        L.pmo_create.argtypes = [C.POINTER(abi.c3d_kernel_input), C.c_int]
        L.pmo_create.restype = C.c_void_p
        L.pmo_destroy.argtypes = [C.c_void_p]
        L.pmo_last_error.restype = C.c_char_p
        L.pmo_run.argtypes = [C.c_void_p, u32p, C.c_uint32, vsp, C.POINTER(abi.c3d_run_params),
                              C.POINTER(abi.c3d_depth_snapshot)]
        L.pmo_evaluate.argtypes = [C.c_void_p, C.c_uint32, u32p, C.c_uint32, vsp, C.c_int,
                                   C.POINTER(abi.c3d_depth_snapshot), C.POINTER(C.POINTER(C.c_float))]
        L.pmo_filter.argtypes = [C.c_void_p, u32p, C.c_uint32, vsp, C.POINTER(abi.c3d_depth_snapshot),
                                 C.POINTER(C.POINTER(C.c_uint8))]
        L.pmo_upsample.argtypes = [C.c_void_p, C.c_uint32, u32p, C.c_uint32, vsp, C.c_float, C.c_float, vsp]
        L.pmo_median.argtypes = [C.c_void_p, C.c_uint32, u32p, C.c_uint32, vsp, vsp]
        L.pmo_densify.argtypes = [C.c_void_p, u32p, C.c_uint32, C.POINTER(abi.c3d_schedule), vsp,
                                  C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.POINTER(C.c_float))]
        L.pmo_run_rows.argtypes = [C.c_void_p, C.c_uint32, vsp, C.POINTER(abi.c3d_run_params), C.c_uint32, C.c_uint32]
        L.pmo_ncc_cost_f32.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, fp]
        L.pmo_ncc_cost_f32.restype = C.c_float
        L.pmo_ncc_cost_f64.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, dp]
        L.pmo_ncc_cost_f64.restype = C.c_double
        from cumulus3d_amd import fusion as _fusion

        L.pmo_fuse_consistency.argtypes = [C.POINTER(_fusion.c3d_fusion_view), C.c_uint32, C.POINTER(_fusion.c3d_fusion_params),
                                           C.POINTER(_fusion.c3d_point_cloud)]
        L.pmo_fuse_consistency.restype = C.c_int
        L.pmo_remove_speckles.argtypes = [fp, fp, fp, C.c_uint32, C.c_uint32, C.POINTER(abi.c3d_post_params), C.POINTER(C.c_uint64)]
        L.pmo_fill_holes.argtypes = [fp, fp, fp, fp, C.POINTER(C.c_uint8), dp, C.c_uint32, C.c_uint32, C.POINTER(abi.c3d_post_params),
                                     C.POINTER(C.c_uint64)]
        from cumulus3d_amd.neighbors import c3d_sparse_scene

        L.pmo_pair_stats.argtypes = [C.POINTER(c3d_sparse_scene), C.POINTER(C.c_uint32), dp]
        L.pmo_depth_ranges.argtypes = [C.POINTER(c3d_sparse_scene), dp, dp, C.POINTER(C.c_uint8)]
        for n in ("pmo_run", "pmo_evaluate", "pmo_filter", "pmo_upsample", "pmo_median", "pmo_densify", "pmo_run_rows",
                  "pmo_remove_speckles", "pmo_fill_holes", "pmo_pair_stats", "pmo_depth_ranges"):
            getattr(L, n).restype = C.c_int
        ip = C.POINTER(C.c_int)
        L.pmo_region_pick.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip]
        L.pmo_region_pick.restype = C.c_int
        L.pmo_transfer_plane.argtypes = [fp, fp, fp, fp]
        L.pmo_transfer_plane.restype = C.c_int
        L.pmo_select_views.argtypes = [C.POINTER(SelectParams), C.c_int, C.c_int, fp, fp, fp, fp, fp, fp, ip, ip]
        L.pmo_select_views.restype = None
        L.pmo_refine_accept.argtypes = [C.POINTER(SelectParams), C.c_int, C.c_int, fp, C.c_float, fp, fp, fp]
        L.pmo_refine_accept.restype = C.c_int
        L.pmo_perturb_normal_det.argtypes = [fp, C.c_float, fp, fp, C.c_float, fp]
        L.pmo_perturb_normal_det.restype = None
        L.pmo_topk_mean.argtypes = [fp, C.c_int, C.c_int]
        L.pmo_topk_mean.restype = C.c_float
        L.pmo_readout_source.argtypes = [C.POINTER(ReadoutParams), fp, C.c_float, fp, fp, C.c_int, C.c_float, C.c_float]
        L.pmo_readout_source.restype = C.c_int
        L.pmo_count_need.argtypes = [C.c_uint32, C.c_uint32]
        L.pmo_count_need.restype = C.c_uint32
        return L


class SelectParams(C.Structure):
    """pmo_select_params: view-selection constants of one half step (tau(t), tau1, 1/(2 beta^2),
    prev_view_weight, lambda, n1, n2)."""
    _fields_ = [("tau_good", C.c_float), ("tau_bad", C.c_float), ("inv2beta2", C.c_float), ("prev_w", C.c_float),
                ("lam", C.c_float), ("n1", C.c_int32), ("n2", C.c_int32)]


class ReadoutParams(C.Structure):
    """pmo_readout_params: read-out constants (cos of the minimum triangulation angle, emission
    normalisation and 1/(2 sigma^2), q_min, maximum geometric error)."""
    _fields_ = [("cos_filter_tri", C.c_float), ("ncc_norm", C.c_float), ("inv2ncc", C.c_float), ("q_min", C.c_float),
                ("filter_gmax", C.c_float)]


def _f32(a, n=None):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def region_pick(cost: np.ndarray, x: int, y: int, region: int):
    """Neighbour chosen for propagation region `region` at (x, y): (qx, qy) or None."""
    c, cp = _f32(cost)
    h, w = c.shape
    qx, qy = C.c_int(-1), C.c_int(-1)
    ok = lib().pmo_region_pick(cp, w, h, x, y, region, C.byref(qx), C.byref(qy))
    return (qx.value, qy.value) if ok else None


def transfer_plane(q, ray_q, ray):
    """The neighbour's plane q = (nx, ny, nz, d) moved onto `ray`, or None."""
    q, qp = _f32(q)
    rq, rqp = _f32(ray_q)
    r, rp = _f32(ray)
    out = np.zeros(4, np.float32)
    ok = lib().pmo_transfer_plane(qp, rqp, rp, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out if ok else None


def select_views(params: SelectParams, prior, m, g=None, vprev: int = 255):
    """View selection over a cost table: prior (nsrc), m and g (nsrc, 9). Returns (wv, den, agg,
    best hypothesis, best view)."""
    pr, prp = _f32(prior)
    n = pr.size
    mm, mp = _f32(np.asarray(m, np.float32).reshape(n, 9))
    gp = None
    if g is not None:
        gg, gp = _f32(np.asarray(g, np.float32).reshape(n, 9))
    wv = np.zeros(max(n, 1), np.float32)
    agg = np.zeros(9, np.float32)
    den = C.c_float()
    best, view = C.c_int(), C.c_int()
    fp = C.POINTER(C.c_float)
    lib().pmo_select_views(C.byref(params), n, vprev, prp, mp, gp, wv.ctypes.data_as(fp), C.cast(C.byref(den), fp),
                           agg.ctypes.data_as(fp), C.byref(best), C.byref(view))
    return wv[:n], float(den.value), agg, best.value, view.value


def refine_accept(params: SelectParams, wv, den: float, m, g=None, bcost: float = 2.0):
    """Refinement acceptance over candidate costs m (ncand, nsrc). Returns (taken index or -1,
    final cost)."""
    w, wp = _f32(wv)
    n = w.size
    mm = np.asarray(m, np.float32).reshape(-1, n)
    mm, mp = _f32(mm)
    gp = None
    if g is not None:
        gg, gp = _f32(np.asarray(g, np.float32).reshape(mm.shape))
    b = C.c_float(bcost)
    k = lib().pmo_refine_accept(C.byref(params), mm.shape[0], n, wp, den, mp, gp, C.cast(C.byref(b), C.POINTER(C.c_float)))
    return k, float(b.value)


def perturb_normal_det(a, u: float, n, ray, phi: float) -> np.ndarray:
    """Normal perturbation given its draws (axis vector a, angle fraction u)."""
    a, ap = _f32(a)
    nn, np_ = _f32(n)
    r, rp = _f32(ray)
    out = np.zeros(3, np.float32)
    lib().pmo_perturb_normal_det(ap, u, np_, rp, phi, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def topk_mean(costs, k: int) -> float:
    c, cp = _f32(costs if len(costs) else [0.0])
    return float(lib().pmo_topk_mean(cp, len(costs), k))


def readout_source(params: ReadoutParams, X, c, n, hok: bool, m: float, geo: float) -> bool:
    """One source's read-out test at the 3-D point X (reference camera at the origin)."""
    x, xp = _f32(X)
    lx = float(np.sqrt(np.float32(x[0]) ** 2 + np.float32(x[1]) ** 2 + np.float32(x[2]) ** 2))
    cc, cp = _f32(c)
    nn, np_ = _f32(n)
    return bool(lib().pmo_readout_source(C.byref(params), xp, lx, cp, np_, int(hok), m, geo))


def count_need(min_num_consistent: int, nsrc: int) -> int:
    return int(lib().pmo_count_need(min_num_consistent, nsrc))


def rcp_sample(z: np.ndarray) -> np.ndarray:
    """The device's per-sample reciprocal of float32 values."""
    z = np.ascontiguousarray(z, np.float32)
    r = np.empty_like(z)
    fp = C.POINTER(C.c_float)
    lib().pmo_rcp_sample(z.ctypes.data_as(fp), r.ctypes.data_as(fp), z.size)
    return r


def build_pyramid(gray: np.ndarray, k, levels: int):
    """Gray pyramid (upsample.rs:78-102), coarsest first: list of (u8 image, [fx, fy, cx, cy])."""
    gray = np.ascontiguousarray(gray, np.uint8)
    h, w = gray.shape
    outs = [np.zeros((h >> l, w >> l), np.uint8) for l in range(1, levels)]
    ko = (C.c_float * (4 * levels))()
    arr = (C.POINTER(C.c_uint8) * max(1, levels - 1))(*[o.ctypes.data_as(C.POINTER(C.c_uint8)) for o in outs])
    lib().pmo_build_pyramid(gray.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, (C.c_double * 4)(*[float(q) for q in k]),
                            levels, arr, ko)
    fine_first = [(gray, [float(ko[j]) for j in range(4)])] + [(outs[l - 1], [float(ko[4 * l + j]) for j in range(4)])
                                                               for l in range(1, levels)]
    return fine_first[::-1]


def fuse_consistency(views_arr, params, capacity: int):
    """Agreement fusion (oracle/fusion_oracle.c): a fusion.Cloud in the reference's point order."""
    from cumulus3d_amd import fusion as _fusion

    pc, arrs = _fusion._cloud(capacity)
    p = params.c()
    rc = lib().pmo_fuse_consistency(views_arr, len(views_arr), C.byref(p), C.byref(pc))
    if rc:
        raise OracleError(f"fusion capacity {capacity} < {pc.count}")
    m = int(pc.count)
    return _fusion.Cloud(arrs[0][:m], arrs[1][:m], arrs[2][:m], arrs[3][:m])


def pair_stats(scene):
    """PairStats::new (oracle/scene_oracle.c; neighbors.rs:47-80): (count, q75), n x n."""
    from cumulus3d_amd.neighbors import SparseScene

    s = scene if isinstance(scene, SparseScene) else SparseScene(scene)
    count = np.zeros((s.n, s.n), np.uint32)
    q75 = np.zeros((s.n, s.n), np.float64)
    lib().pmo_pair_stats(C.byref(s.c), count.ctypes.data_as(C.POINTER(C.c_uint32)), q75.ctypes.data_as(C.POINTER(C.c_double)))
    return count, q75


def select(count, q75, r, max_views, min_angle_rad):
    """PairStats::select (neighbors.rs:83-103), restated."""
    cands = [(int(count[r, o]), o) for o in range(count.shape[0])
             if o != r and count[r, o] > 0 and not (min_angle_rad > 0.0 and q75[r, o] < min_angle_rad)]
    cands.sort(key=lambda c: (-c[0], c[1]))
    return [o for _, o in cands[:max_views]]


def select_sources(scene, max_views=10, min_angle_deg=1.0):
    """select_all (neighbors.rs:154-157) over the oracle's pair statistics."""
    import math

    count, q75 = pair_stats(scene)
    return [select(count, q75, r, max_views, math.radians(min_angle_deg)) for r in range(count.shape[0])]


def depth_ranges(scene):
    """depth_ranges (oracle/scene_oracle.c; neighbors.rs:162-189): per view (lo, hi) or None."""
    from cumulus3d_amd.neighbors import SparseScene

    s = scene if isinstance(scene, SparseScene) else SparseScene(scene)
    lo = np.zeros(s.n, np.float64)
    hi = np.zeros(s.n, np.float64)
    has = np.zeros(s.n, np.uint8)
    lib().pmo_depth_ranges(C.byref(s.c), lo.ctypes.data_as(C.POINTER(C.c_double)), hi.ctypes.data_as(C.POINTER(C.c_double)),
                           has.ctypes.data_as(C.POINTER(C.c_uint8)))
    return [(float(lo[v]), float(hi[v])) if has[v] else None for v in range(s.n)]


def remove_speckles(depth, normal, cost, params):
    """remove_speckles (oracle/post_oracle.c; postproc.rs:77-133) in place; returns the count."""
    h, w = depth.shape
    n = C.c_uint64()
    p = params.c()
    fp = C.POINTER(C.c_float)
    lib().pmo_remove_speckles(depth.ctypes.data_as(fp), normal.ctypes.data_as(fp), cost.ctypes.data_as(fp), w, h, C.byref(p),
                              C.byref(n))
    return int(n.value)


def fill_holes(depth, normal, cost, raw_cost, gray, k, params):
    """fill_holes (oracle/post_oracle.c; postproc.rs:135-204) in place; returns the count."""
    h, w = depth.shape
    n = C.c_uint64()
    p = params.c()
    fp = C.POINTER(C.c_float)
    gray = np.ascontiguousarray(gray, np.uint8)
    lib().pmo_fill_holes(depth.ctypes.data_as(fp), normal.ctypes.data_as(fp), cost.ctypes.data_as(fp), raw_cost.ctypes.data_as(fp),
                         gray.ctypes.data_as(C.POINTER(C.c_uint8)), (C.c_double * 4)(*[float(q) for q in k]), w, h, C.byref(p),
                         C.byref(n))
    return int(n.value)


class OracleError(RuntimeError):
    pass


def _check(rc: int):
    if rc != 0:
        raise OracleError(f"oracle error {rc}: {lib().pmo_last_error().decode()}")


class Oracle:
    """CPU restatement of a PatchStereo session over a KernelInput (same operations as the C ABI)."""

    def __init__(self, inp: abi.KernelInput, threads: Optional[int] = None, rcp_mode: int = 1):
        self.inp = inp
        L = lib()
        L.pmo_set_rcp_mode(rcp_mode)
        self.threads = threads or os.cpu_count() or 1
        self.ctx = L.pmo_create(C.byref(inp.c), self.threads)
        if not self.ctx:
            raise OracleError(L.pmo_last_error().decode())

    def close(self):
        if self.ctx:
            lib().pmo_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def run(self, views: Sequence[int], states: Sequence[abi.ViewState], params: abi.c3d_run_params,
            snapshot: Optional[abi.DepthSnapshot] = None):
        arr = abi.states_array(states)
        _check(lib().pmo_run(self.ctx, abi.views_array(views), len(views), arr, C.byref(params),
                             C.byref(snapshot.c) if snapshot else None))

    def evaluate(self, level: int, views, states, geometric=False, snapshot=None):
        outs = [np.zeros((s.height, s.width), np.float32) for s in states]
        _check(lib().pmo_evaluate(self.ctx, level, abi.views_array(views), len(views), abi.states_array(states),
                                  int(geometric), C.byref(snapshot.c) if snapshot else None,
                                  abi.ptr_array(outs, C.c_float)))
        return outs

    def filter(self, views, states, snapshot):
        outs = [np.zeros((s.height, s.width), np.uint8) for s in states]
        _check(lib().pmo_filter(self.ctx, abi.views_array(views), len(views), abi.states_array(states),
                                C.byref(snapshot.c), abi.ptr_array(outs, C.c_uint8)))
        return outs

    def upsample(self, level, views, low, sigma_s, sigma_c):
        outs = []
        for v in views:
            h, w = self.inp.level_shape(v, level)
            outs.append(abi.ViewState(w, h))
        _check(lib().pmo_upsample(self.ctx, level, abi.views_array(views), len(views), abi.states_array(low),
                                  sigma_s, sigma_c, abi.states_array(outs)))
        return outs

    def median_filter(self, level, views, states):
        outs = [abi.ViewState(s.width, s.height) for s in states]
        _check(lib().pmo_median(self.ctx, level, abi.views_array(views), len(views), abi.states_array(states),
                                abi.states_array(outs)))
        for o, s in zip(outs, states):
            o.cost[...] = s.cost
        return outs

    def densify(self, views, schedule: abi.Schedule):
        """compute_depth_maps over every scale (densify.rs:350-454): returns top-scale (states with
        thresholded depth, counts, raw depth)."""
        sch = schedule.c([self.inp.view_k[v] for v in views] if schedule.post is not None else None)
        top = self.inp.num_levels - 1
        outs, counts, raws = [], [], []
        for v in views:
            h, w = self.inp.level_shape(v, top)
            outs.append(abi.ViewState(w, h))
            counts.append(np.zeros((h, w), np.uint8))
            raws.append(np.zeros((h, w), np.float32))
        _check(lib().pmo_densify(self.ctx, abi.views_array(views), len(views), C.byref(sch), abi.states_array(outs),
                                 abi.ptr_array(counts, C.c_uint8), abi.ptr_array(raws, C.c_float)))
        return outs, counts, raws

    def run_rows(self, view, state, params, row0, row1):
        arr = abi.states_array([state])
        _check(lib().pmo_run_rows(self.ctx, view, arr, C.byref(params), row0, row1))

    def ncc_cost_f32(self, level, view, slot, x, y, plane):
        p = (C.c_float * 4)(*[float(q) for q in plane])
        return lib().pmo_ncc_cost_f32(self.ctx, level, view, slot, x, y, p)

    def ncc_cost_f64(self, level, view, slot, x, y, plane):
        p = (C.c_double * 4)(*[float(q) for q in plane])
        return lib().pmo_ncc_cost_f64(self.ctx, level, view, slot, x, y, p)
