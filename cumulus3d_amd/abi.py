"""ctypes mirror of include/c3d_pm.h (the C ABI of the PatchStereo engine).

The structures follow the reference's boundary types (crates/dense/src/kernel.rs:81-189):
``LevelImage``, ``PairGeometry``, ``KernelView``, ``KernelInput``, ``ViewState``,
``DepthSnapshot``, ``RunParams``; ``PmParams``/``FilterParams`` defaults are those of
crates/dense/src/params.rs:126-154,196-207 with the fast-profile tuning of params.rs:452-459
available through :func:`fast_profile`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields
from typing import List, Optional, Sequence

import numpy as np

MAX_SOURCES = 32
# the half step's compiled workgroup shapes (rows of active pixels, lanes per pixel), in the order of
# c3d_pm_autotune's timings (C3D_TILE_SHAPES)
TILE_SHAPES = ((1, 1), (2, 1), (4, 1), (4, 2))

C3D_OK = 0
C3D_INVALID_ARGUMENT = 1
C3D_DEVICE = 2
C3D_NOT_AVAILABLE = 3
C3D_COMPILE = 4


class c3d_pm_options(C.Structure):
    _fields_ = [
        ("batch_pixels", C.c_int64),
        ("max_batch", C.c_int32),
        ("tile_rows", C.c_int32),
        ("threads_per_pixel", C.c_int32),
        ("use_graph", C.c_int32),
        ("any_radius", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class c3d_level_image(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("gray", C.POINTER(C.c_uint8)),
        ("k", C.c_float * 4),
    ]


class c3d_pair_geometry(C.Structure):
    _fields_ = [("r", C.c_float * 9), ("t", C.c_float * 3), ("center", C.c_float * 3)]


class c3d_kernel_view(C.Structure):
    _fields_ = [
        ("key", C.c_uint64),
        ("levels", C.POINTER(c3d_level_image)),
        ("num_sources", C.c_uint32),
        ("pad0", C.c_uint32),
        ("sources", C.POINTER(C.c_uint32)),
        ("pairs", C.POINTER(c3d_pair_geometry)),
        ("depth_min", C.c_float),
        ("depth_max", C.c_float),
    ]


class c3d_pm_params(C.Structure):
    _fields_ = [
        ("window_radius", C.c_int32),
        ("window_step", C.c_int32),
        ("sigma_spatial", C.c_float),
        ("sigma_color", C.c_float),
        ("tau0", C.c_float),
        ("tau_alpha", C.c_float),
        ("tau1", C.c_float),
        ("beta", C.c_float),
        ("n1", C.c_uint32),
        ("n2", C.c_uint32),
        ("prev_view_weight", C.c_float),
        ("min_triangulation_angle_deg", C.c_float),
        ("incident_angle_sigma", C.c_float),
        ("geom_lambda", C.c_float),
        ("geom_max_cost", C.c_float),
        ("top_k", C.c_uint32),
        ("eps0_photometric", C.c_float),
        ("eps0_geometric", C.c_float),
        ("phi0_photometric_deg", C.c_float),
        ("phi0_geometric_deg", C.c_float),
        ("eps_min", C.c_float),
        ("phi_min_deg", C.c_float),
        ("weak_texture_var", C.c_float),
        ("pad0", C.c_int32),
    ]


class c3d_filter_params(C.Structure):
    _fields_ = [
        ("min_ncc", C.c_float),
        ("min_triangulation_angle_deg", C.c_float),
        ("min_num_consistent", C.c_uint32),
        ("geom_max_cost", C.c_float),
        ("ncc_sigma", C.c_float),
        ("median_filter", C.c_int32),
    ]


class c3d_kernel_input(C.Structure):
    _fields_ = [
        ("views", C.POINTER(c3d_kernel_view)),
        ("num_views", C.c_uint32),
        ("num_levels", C.c_uint32),
        ("pm", c3d_pm_params),
        ("filter", c3d_filter_params),
        ("seed", C.c_uint64),
    ]


class c3d_view_state(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("depth", C.POINTER(C.c_float)),
        ("normal", C.POINTER(C.c_float)),
        ("cost", C.POINTER(C.c_float)),
        ("prior", C.POINTER(C.c_float)),
    ]


class c3d_depth_snapshot(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("level", C.c_uint32),
        ("num_maps", C.c_uint32),
        ("maps", C.POINTER(C.POINTER(C.c_float))),
    ]


class c3d_run_params(C.Structure):
    _fields_ = [
        ("level", C.c_uint32),
        ("geometric", C.c_int32),
        ("random_init", C.c_int32),
        ("iterations", C.c_uint32),
        ("run_id", C.c_uint32),
        ("use_prior", C.c_int32),
    ]


class c3d_post_params(C.Structure):
    _fields_ = [
        ("remove_speckles", C.c_int32),
        ("speckle_rel_depth", C.c_float),
        ("speckle_min_area_frac", C.c_double),
        ("speckle_keep_cost", C.c_float),
        ("fill_holes", C.c_int32),
        ("fill_max_frac", C.c_double),
        ("fill_edge_step", C.c_float),
        ("fill_max_cost", C.c_float),
    ]


class c3d_schedule(C.Structure):
    _fields_ = [
        ("photometric_iters", C.c_uint32),
        ("photometric_iters_upper", C.c_uint32),
        ("geometric_rounds", C.c_uint32),
        ("geometric_iters", C.c_uint32),
        ("first_run_id", C.c_uint32),
        ("restorer_random_init", C.c_int32),
        ("jbu_sigma_spatial", C.c_float),
        ("jbu_sigma_color", C.c_float),
        ("restorer_threshold", C.c_float),
        ("pad0", C.c_uint32),
        ("post", c3d_post_params),
        ("post_k", C.POINTER(C.c_double)),
    ]


class c3d_timing(C.Structure):
    _fields_ = [
        ("photometric_ms", C.c_double),
        ("geometric_ms", C.c_double),
        ("filter_ms", C.c_double),
        ("total_ms", C.c_double),
        ("half_kernel_ms", C.c_double),
        ("half_launches", C.c_uint32),
        ("pad0", C.c_uint32),
        ("init_ms", C.c_double),
        ("post_ms", C.c_double),
        ("post_removed", C.c_uint64),
        ("post_filled", C.c_uint64),
    ]


# ---------------------------------------------------------------- Python-side parameter records


@dataclass
class PostParams:
    """PostParams (crates/dense/src/postproc.rs:40-73), the reference's defaults; `fast()` adds the
    fast profile's hole filling (params.rs:456)."""

    remove_speckles: bool = True
    speckle_rel_depth: float = 0.02
    speckle_min_area_frac: float = 2e-4
    speckle_keep_cost: float = 0.2
    fill_holes: bool = False
    fill_max_frac: float = 0.01
    fill_edge_step: float = 0.06
    fill_max_cost: float = 1.0

    @classmethod
    def fast(cls) -> "PostParams":
        return cls(fill_holes=True)

    def c(self) -> c3d_post_params:
        return c3d_post_params(int(self.remove_speckles), self.speckle_rel_depth, self.speckle_min_area_frac, self.speckle_keep_cost,
                               int(self.fill_holes), self.fill_max_frac, self.fill_edge_step, self.fill_max_cost)


def post_k_array(ks) -> "C.Array":
    """4 doubles (fx, fy, cx, cy) per view, for c3d_schedule.post_k."""
    flat = [float(q) for k in ks for q in k]
    return (C.c_double * max(len(flat), 1))(*flat)


@dataclass
class Schedule:
    """Per-scale iteration schedule (LevelSchedule, crates/dense/src/params.rs:476-506) and the
    densify options it reads (jbu sigmas, restorer threshold: params.rs:423-426). The defaults are
    the fast profile; `quality()` is MvsProfile::Quality."""

    photometric_iters: int = 6
    photometric_iters_upper: int = 3
    geometric_rounds: int = 2
    geometric_iters: int = 2
    restorer_random_init: bool = False
    jbu_sigma_spatial: float = 1.0
    jbu_sigma_color: float = 0.04
    restorer_threshold: float = 0.1
    first_run_id: int = 1
    post: Optional[PostParams] = None  # read-out post-processing (densify.rs:455-462); None = off

    @classmethod
    def fast(cls) -> "Schedule":
        return cls()

    @classmethod
    def quality(cls) -> "Schedule":
        return cls(photometric_iters=7, photometric_iters_upper=6, geometric_rounds=2, geometric_iters=6,
                   restorer_random_init=True)

    def c(self, view_k=None) -> c3d_schedule:
        """The C struct; `view_k` = top-scale (fx, fy, cx, cy) of each densified view, in order
        (needed when post.fill_holes)."""
        sch = c3d_schedule(self.photometric_iters, self.photometric_iters_upper, self.geometric_rounds, self.geometric_iters,
                           self.first_run_id, int(self.restorer_random_init), self.jbu_sigma_spatial, self.jbu_sigma_color,
                           self.restorer_threshold, 0)
        if self.post is not None:
            sch.post = self.post.c()
        if view_k is not None:
            sch._k = post_k_array(view_k)
            sch.post_k = C.cast(sch._k, C.POINTER(C.c_double))
        return sch


@dataclass
class PmParams:
    """PmParams (params.rs:76-154), defaults as in the reference."""

    window_radius: int = 5
    window_step: int = 2
    sigma_spatial: float = -1.0
    sigma_color: float = 0.2
    tau0: float = 0.8
    tau_alpha: float = 90.0
    tau1: float = 1.2
    beta: float = 0.3
    n1: int = 2
    n2: int = 3
    prev_view_weight: float = 0.2
    min_triangulation_angle_deg: float = 1.0
    incident_angle_sigma: float = 0.9
    geom_lambda: float = 0.2
    geom_max_cost: float = 3.0
    top_k: int = 3
    eps0_photometric: float = 0.2
    eps0_geometric: float = 0.05
    phi0_photometric_deg: float = 30.0
    phi0_geometric_deg: float = 10.0
    eps_min: float = 0.005
    phi_min_deg: float = 1.0
    weak_texture_var: float = 0.0

    def sigma_s(self) -> float:
        """PmParams::sigma_s (params.rs:158-164)."""
        return self.sigma_spatial if self.sigma_spatial > 0.0 else float(self.window_radius)

    def to_c(self) -> c3d_pm_params:
        c = c3d_pm_params()
        for f in fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c


@dataclass
class FilterParams:
    """FilterParams (params.rs:179-207), defaults as in the reference."""

    min_ncc: float = 0.1
    min_triangulation_angle_deg: float = 3.0
    min_num_consistent: int = 2
    geom_max_cost: float = 1.0
    ncc_sigma: float = 0.6
    median_filter: bool = True

    def to_c(self) -> c3d_filter_params:
        c = c3d_filter_params()
        for f in fields(self):
            setattr(c, f.name, int(getattr(self, f.name)) if f.name == "median_filter" else getattr(self, f.name))
        return c


DEFAULT_SEED = 0x005E_ED0F_DE75  # DensifyOptions seed (params.rs:427)


def fast_profile() -> tuple[PmParams, FilterParams]:
    """Fast profile (params.rs:440-459): window step 2, tuned filter (min_ncc 0.2, geometric
    tolerance 0.5 px, 3 consistent views, no median filter)."""
    pm = PmParams(window_step=2)
    fl = FilterParams(min_ncc=0.2, geom_max_cost=0.5, min_num_consistent=3, median_filter=False)
    return pm, fl


def quality_profile() -> tuple[PmParams, FilterParams]:
    """Quality profile (params.rs:440-451): window step 1, untuned filter."""
    return PmParams(window_step=1), FilterParams()


# ---------------------------------------------------------------- input builder


@dataclass
class LevelData:
    """One view image at one scale: u8 gray (H, W) and k = (fx, fy, cx, cy), integer-centre."""

    gray: np.ndarray
    k: Sequence[float]


@dataclass
class ViewData:
    key: int
    levels: List[LevelData]
    sources: List[int]
    pairs: List[tuple]  # (r[9], t[3], center[3]) float
    depth_min: float
    depth_max: float
    k: Optional[List[float]] = None  # the image's own fx, fy, cx, cy in double (fill_holes reads view.k)


class KernelInput:
    """Owns the ctypes image of a ``KernelInput`` (kernel.rs:122-135) and every buffer it points to."""

    def __init__(self, views: Sequence[ViewData], pm: PmParams, filt: FilterParams, seed: int = DEFAULT_SEED):
        self._keep: list = []
        nl = len(views[0].levels) if views else 0
        # top-scale intrinsics in double per view (the read-out post-processing's view.k)
        self.view_k = [[float(q) for q in (v.k if v.k is not None else v.levels[-1].k)] for v in views]
        carr = (c3d_kernel_view * max(len(views), 1))()
        for i, v in enumerate(views):
            if len(v.levels) != nl:
                raise ValueError("every view needs the same number of levels")
            lv = (c3d_level_image * nl)()
            for l, L in enumerate(v.levels):
                g = np.ascontiguousarray(L.gray, dtype=np.uint8)
                self._keep.append(g)
                lv[l].width = g.shape[1]
                lv[l].height = g.shape[0]
                lv[l].gray = g.ctypes.data_as(C.POINTER(C.c_uint8))
                for q in range(4):
                    lv[l].k[q] = float(L.k[q])
            src = (C.c_uint32 * max(len(v.sources), 1))(*v.sources)
            prs = (c3d_pair_geometry * max(len(v.pairs), 1))()
            for j, (r, t, c) in enumerate(v.pairs):
                for q in range(9):
                    prs[j].r[q] = float(r[q])
                for q in range(3):
                    prs[j].t[q] = float(t[q])
                    prs[j].center[q] = float(c[q])
            self._keep += [lv, src, prs]
            carr[i].key = v.key & 0xFFFFFFFFFFFFFFFF
            carr[i].levels = lv
            carr[i].num_sources = len(v.sources)
            carr[i].sources = src
            carr[i].pairs = prs
            carr[i].depth_min = v.depth_min
            carr[i].depth_max = v.depth_max
        self._keep.append(carr)
        self.c = c3d_kernel_input()
        self.c.views = carr
        self.c.num_views = len(views)
        self.c.num_levels = nl
        self.c.pm = pm.to_c()
        self.c.filter = filt.to_c()
        self.c.seed = seed & 0xFFFFFFFFFFFFFFFF
        self.views = list(views)
        self.pm = pm
        self.filter = filt
        self.seed = seed
        self.num_levels = nl

    def level_shape(self, view: int, level: int) -> tuple[int, int]:
        g = self.views[view].levels[level].gray
        return g.shape[0], g.shape[1]


class ViewState:
    """ViewState (kernel.rs:137-160): depth (H, W), normal (H, W, 3), cost (H, W), optional prior (H, W, 4)."""

    def __init__(self, width: int, height: int, prior: bool = False):
        self.width = width
        self.height = height
        self.depth = np.zeros((height, width), np.float32)
        self.normal = np.zeros((height, width, 3), np.float32)
        self.cost = np.full((height, width), 2.0, np.float32)
        self.prior = np.zeros((height, width, 4), np.float32) if prior else None

    def copy(self) -> "ViewState":
        s = ViewState(self.width, self.height)
        s.depth = self.depth.copy()
        s.normal = self.normal.copy()
        s.cost = self.cost.copy()
        s.prior = None if self.prior is None else self.prior.copy()
        return s

    def to_c(self) -> c3d_view_state:
        for a in (self.depth, self.normal, self.cost):
            assert a.flags.c_contiguous and a.dtype == np.float32
        c = c3d_view_state()
        c.width = self.width
        c.height = self.height
        fp = C.POINTER(C.c_float)
        c.depth = self.depth.ctypes.data_as(fp)
        c.normal = self.normal.ctypes.data_as(fp)
        c.cost = self.cost.ctypes.data_as(fp)
        c.prior = self.prior.ctypes.data_as(fp) if self.prior is not None else None
        return c


def states_array(states: Sequence[ViewState]):
    arr = (c3d_view_state * max(len(states), 1))()
    for i, s in enumerate(states):
        arr[i] = s.to_c()
    return arr


def views_array(views: Sequence[int]):
    return (C.c_uint32 * max(len(views), 1))(*views)


class DepthSnapshot:
    """DepthSnapshot (kernel.rs:162-171); maps[v] is a (H, W) float32 array or None."""

    def __init__(self, id: int, level: int, maps: Sequence[Optional[np.ndarray]]):
        self.maps = [None if m is None else np.ascontiguousarray(m, np.float32) for m in maps]
        ptrs = (C.POINTER(C.c_float) * max(len(self.maps), 1))()
        for i, m in enumerate(self.maps):
            ptrs[i] = m.ctypes.data_as(C.POINTER(C.c_float)) if m is not None else None
        self._ptrs = ptrs
        self.c = c3d_depth_snapshot()
        self.c.id = id
        self.c.level = level
        self.c.num_maps = len(self.maps)
        self.c.maps = ptrs


def run_params(level=0, geometric=False, random_init=False, iterations=0, run_id=0, use_prior=False) -> c3d_run_params:
    p = c3d_run_params()
    p.level = level
    p.geometric = int(geometric)
    p.random_init = int(random_init)
    p.iterations = iterations
    p.run_id = run_id
    p.use_prior = int(use_prior)
    return p


def ptr_array(arrays, ctype):
    arr = (C.POINTER(ctype) * max(len(arrays), 1))()
    for i, a in enumerate(arrays):
        arr[i] = a.ctypes.data_as(C.POINTER(ctype))
    return arr
