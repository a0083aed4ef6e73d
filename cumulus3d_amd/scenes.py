"""Scenes and kernel inputs for the PatchStereo engine.

* :class:`DenseView` / :class:`DenseScene`: the reference's scene types (crates/dense/src/scene.rs:57-160):
  integer-centre intrinsics, world->camera pose, 8-bit gray.
* :func:`pair_geometry`: reference->source relative geometry (crates/dense/src/densify.rs:194-205).
* :func:`make_synth_scene`: the reference's accuracy fixture, a textured floor with two boxes seen by a
  3x3 camera grid with ground-truth depth and normals (crates/dense/src/synthetic.rs:36-292).
* sparse points (ScenePoint) for the neighbour selection and depth ranges of neighbors.py
  (crates/dense/src/neighbors.rs:39-189).
* :func:`make_terrain_scene`: seeded fBm terrain imaged by oblique pinhole cameras (no reference
  counterpart; rendered by csrc/terrain.c).
* :func:`kernel_input`: builds the C-ABI ``KernelInput`` (key = FNV-1a of the image name,
  densify.rs:295-309).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import abi

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

M64 = 0xFFFFFFFFFFFFFFFF


# ---------------------------------------------------------------- integer hashes (math.rs:31-47)


def hash_bytes(b: bytes) -> int:
    """FNV-1a 64 (crates/dense/src/math.rs:31-39)."""
    h = 0xCBF29CE484222325
    for x in b:
        h ^= x
        h = (h * 0x00000100000001B3) & M64
    return h


def mix64(z):
    """splitmix64 finaliser (math.rs:41-47); works on Python ints and uint64 arrays."""
    if isinstance(z, np.ndarray):
        with np.errstate(over="ignore"):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


# ---------------------------------------------------------------- scene types


@dataclass
class DenseView:
    """One densification view (scene.rs:57-142). k = (fx, fy, cx, cy) integer-centre; r, t world->camera."""

    name: str
    width: int
    height: int
    k: np.ndarray
    r: np.ndarray
    t: np.ndarray
    gray: np.ndarray

    def center(self) -> np.ndarray:
        return -self.r.T @ self.t

    def to_cam(self, x):
        return self.r @ np.asarray(x, float) + self.t

    def to_world(self, c):
        return self.r.T @ (np.asarray(c, float) - self.t)


@dataclass
class ScenePoint:
    xyz: np.ndarray
    views: List[int]


@dataclass
class DenseScene:
    views: List[DenseView]
    points: List[ScenePoint] = field(default_factory=list)
    gt_depth: Optional[List[np.ndarray]] = None
    gt_normal: Optional[List[np.ndarray]] = None


def pair_geometry(scene: DenseScene, r: int, s: int):
    """Reference->source relative geometry (densify.rs:194-205): R = R_s R_r^T, t = t_s - R t_r,
    c = -R^T t; returned as float32 (r[9], t[3], center[3])."""
    vr, vs = scene.views[r], scene.views[s]
    rr = vs.r @ vr.r.T
    t = vs.t - rr @ vr.t
    c = -(rr.T @ t)
    return (rr.reshape(9).astype(np.float32), t.astype(np.float32), c.astype(np.float32))


# ---------------------------------------------------------------- reference synthetic scene


@dataclass
class SynthBox:
    min: tuple
    max: tuple


@dataclass
class SynthConfig:
    """SynthConfig (synthetic.rs:85-122)."""

    width: int = 384
    height: int = 288
    focal: float = 300.0
    grid: tuple = (3, 3)
    spacing: float = 3.0
    altitude: float = 20.0
    boxes: tuple = (SynthBox((-3.0, -2.0, 0.0), (1.0, 2.5, 4.0)), SynthBox((3.0, 2.0, 0.0), (6.0, 5.0, 2.0)))
    seed: int = 11
    num_points: int = 3000


def _hash2(ix, iy, seed: int):
    """synthetic.rs:52-55 on int64 arrays."""
    with np.errstate(over="ignore"):
        z = (ix.astype(np.int64).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (
            iy.astype(np.int64).astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
        ) ^ np.uint64(seed & M64)
    return (mix64(z) >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def _value_noise(x, y, seed: int):
    """synthetic.rs:57-66."""
    x0, y0 = np.floor(x), np.floor(y)
    s = lambda t: t * t * (3.0 - 2.0 * t)
    sx, sy = s(x - x0), s(y - y0)
    ix, iy = x0.astype(np.int64), y0.astype(np.int64)
    a, b = _hash2(ix, iy, seed), _hash2(ix + 1, iy, seed)
    c, d = _hash2(ix, iy + 1, seed), _hash2(ix + 1, iy + 1, seed)
    top = a + (b - a) * sx
    bot = c + (d - c) * sx
    return top + (bot - top) * sy


def synth_texture(u, v, seed: int):
    """synthetic.rs:68-74."""
    blocks = _hash2(np.floor(u / 0.18).astype(np.int64), np.floor(v / 0.18).astype(np.int64), seed ^ 0x55)
    fine = _value_noise(u / 0.09, v / 0.09, seed)
    coarse = _value_noise(u / 1.5, v / 1.5, seed + 7)
    return np.clip(0.45 * blocks + 0.3 * fine + 0.25 * coarse, 0.0, 1.0)


def _cast(o, d, cfg: SynthConfig):
    """First hit of rays o + t d with the floor and the boxes (synthetic.rs:136-188).
    d: (N, 3). Returns (t, world normal (N,3), texture value, hit mask)."""
    n = d.shape[0]
    best_t = np.full(n, np.inf)
    best_n = np.zeros((n, 3))
    best_v = np.zeros(n)
    hit = np.zeros(n, bool)
    down = d[:, 2] < -1e-12
    t = np.where(down, -o[2] / np.where(down, d[:, 2], -1.0), np.inf)
    ok = down & (t > 0)
    p0 = o[0] + t * d[:, 0]
    p1 = o[1] + t * d[:, 1]
    best_t = np.where(ok, t, best_t)
    best_n[ok] = (0.0, 0.0, 1.0)
    best_v = np.where(ok, synth_texture(np.where(ok, p0, 0.0), np.where(ok, p1, 0.0), cfg.seed), best_v)
    hit |= ok
    for bi, b in enumerate(cfg.boxes):
        t0 = np.zeros(n)
        t1 = np.full(n, np.inf)
        axis = np.zeros(n, np.int64)
        sign = np.zeros(n)
        good = np.ones(n, bool)
        for a in range(3):
            par = np.abs(d[:, a]) < 1e-12
            good &= ~(par & ((o[a] < b.min[a]) | (o[a] > b.max[a])))
            da = np.where(par, 1.0, d[:, a])
            ta = (b.min[a] - o[a]) / da
            tb = (b.max[a] - o[a]) / da
            sw = ta > tb
            s = np.where(sw, 1.0, -1.0)
            ta, tb = np.where(sw, tb, ta), np.where(sw, ta, tb)
            upd = (~par) & (ta > t0)
            t0 = np.where(upd, ta, t0)
            axis = np.where(upd, a, axis)
            sign = np.where(upd, s, sign)
            t1 = np.where(~par, np.minimum(t1, tb), t1)
        good &= ~((t0 > t1) | (t0 <= 0.0))
        good &= ~(hit & (best_t <= t0))
        if not good.any():
            continue
        p = o[None, :] + t0[:, None] * d
        nn = np.zeros((n, 3))
        nn[np.arange(n), axis] = sign
        u = np.where(axis == 0, p[:, 1], p[:, 0])
        v = np.where(axis == 2, p[:, 1], p[:, 2])
        tex = synth_texture(u + 17.0 * bi, v - 5.0 * axis, cfg.seed + 3)
        best_t = np.where(good, t0, best_t)
        best_n[good] = nn[good]
        best_v = np.where(good, tex, best_v)
        hit |= good
    return best_t, best_n, best_v, hit


def _unit(a):
    return a / np.linalg.norm(a)


def _round_half_away(x):
    return np.where(x >= 0, np.floor(x + 0.5), np.ceil(x - 0.5))


def surface_distance(points: np.ndarray, cfg: SynthConfig = SynthConfig()) -> np.ndarray:
    """Distance of world points (N, 3) to the floor-and-boxes surface (SynthScene::surface_distance,
    synthetic.rs:296-310): the floor z = 0, or a box's nearest face from inside, its closest point
    from outside."""
    p = np.asarray(points, np.float64).reshape(-1, 3)
    best = np.abs(p[:, 2])
    for b in cfg.boxes:
        lo, hi = np.asarray(b.min, np.float64), np.asarray(b.max, np.float64)
        inside = np.all((p >= lo) & (p <= hi), axis=1)
        d_in = np.minimum(p - lo, hi - p).min(axis=1)
        q = np.maximum(np.maximum(lo - p, 0.0), p - hi)
        d_out = np.sqrt((q * q).sum(axis=1))
        best = np.minimum(best, np.where(inside, d_in, d_out))
    return best


def make_synth_scene(cfg: SynthConfig = SynthConfig()) -> DenseScene:
    """The reference's floor+boxes scene with ground truth (synthetic.rs:190-292)."""
    w, h, f = cfg.width, cfg.height, cfg.focal
    cx, cy = (w - 1.0) / 2.0, (h - 1.0) / 2.0
    gx, gy = cfg.grid
    views, depths, normals = [], [], []
    ys, xs = np.mgrid[0:h, 0:w]
    xs = xs.reshape(-1).astype(np.float64)
    ys = ys.reshape(-1).astype(np.float64)
    for j in range(gy):
        for i in range(gx):
            c = np.array([(i - (gx - 1) / 2.0) * cfg.spacing, (j - (gy - 1) / 2.0) * cfg.spacing, cfg.altitude])
            target = np.array([c[0] * 0.5, c[1] * 0.5, 0.0])
            zc = _unit(target - c)
            xc = _unit(np.cross(zc, [0.0, 1.0, 0.0]))
            yc = np.cross(zc, xc)
            r = np.stack([xc, yc, zc])
            t = -r @ c
            acc = np.zeros(w * h)
            dm = np.zeros(w * h, np.float32)
            nm = np.zeros((w * h, 3), np.float32)
            for sx, sy in [(-0.25, -0.25), (0.25, -0.25), (-0.25, 0.25), (0.25, 0.25), (0.0, 0.0)]:
                rc = np.stack([(xs + sx - cx) / f, (ys + sy - cy) / f, np.ones_like(xs)], 1)
                dw = rc @ r  # rows: xc*rc0 + yc*rc1 + zc*rc2
                dw /= np.linalg.norm(dw, axis=1, keepdims=True)
                tt, nw, val, hit = _cast(c, dw, cfg)
                if sx == 0.0 and sy == 0.0:
                    p = c[None, :] + np.where(hit, tt, 0.0)[:, None] * dw
                    dm = np.where(hit, (p @ zc) + t[2], 0.0).astype(np.float32)
                    nc = nw @ r.T
                    nm = np.where(hit[:, None], nc, 0.0).astype(np.float32)
                else:
                    acc += np.where(hit, val, 0.0)
            v = np.clip(_round_half_away(acc / 4.0 * 230.0 + 15.0), 0, 255)
            b = np.floor(v * 0.8)
            gray = np.clip(_round_half_away(0.2126 * v + 0.7152 * v + 0.0722 * b), 0, 255).astype(np.uint8)
            views.append(
                DenseView(
                    name=f"synth_{j}_{i}.png",
                    width=w,
                    height=h,
                    k=np.array([f, f, cx, cy]),
                    r=r,
                    t=t,
                    gray=gray.reshape(h, w),
                )
            )
            depths.append(dm.reshape(h, w))
            normals.append(nm.reshape(h, w, 3))
    # sparse points (synthetic.rs:253-290)
    points = []
    s = (cfg.seed ^ 0xABC) & M64

    def rnd():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M64
        return (mix64(s) >> 11) / float(1 << 53)

    nv = len(views)
    for kk in range(cfg.num_points):
        v = kk % nv
        x = int(_round_half_away(np.float64(rnd() * (w - 1))))
        y = int(_round_half_away(np.float64(rnd() * (h - 1))))
        d = float(depths[v][y, x])
        if d <= 0.0:
            continue
        p = views[v].to_world([(x - cx) / f * d, (y - cy) / f * d, d])
        obs = []
        for o, ov in enumerate(views):
            cc = ov.to_cam(p)
            if cc[2] <= 0.0:
                continue
            u, vv = ov.k[0] * cc[0] / cc[2] + ov.k[2], ov.k[1] * cc[1] / cc[2] + ov.k[3]
            ui, vi = _round_half_away(u), _round_half_away(vv)
            if ui < 0 or vi < 0 or ui >= w or vi >= h:
                continue
            dt = float(depths[o][int(vi), int(ui)])
            if abs(dt - cc[2]) < 0.01 * cc[2]:
                obs.append(o)
        if len(obs) >= 2:
            points.append(ScenePoint(np.asarray(p), obs))
    return DenseScene(views, points, depths, normals)


# ---------------------------------------------------------------- terrain scenes


class _TerrainParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("extent", C.c_double),
        ("relief", C.c_double),
        ("base_wavelength", C.c_double),
        ("octaves", C.c_int32),
        ("pad0", C.c_int32),
        ("albedo_scale", C.c_double),
        ("sky", C.c_double),
    ]


class _TerrainCamera(C.Structure):
    _fields_ = [
        ("k", C.c_double * 4),
        ("r", C.c_double * 9),
        ("t", C.c_double * 3),
        ("width", C.c_int32),
        ("height", C.c_int32),
    ]


_scenes_lib = None


def _terrain_lib():
    global _scenes_lib
    if _scenes_lib is None:
        path = os.path.join(_LIB_DIR, "libc3d_scenes.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        lib.c3d_render_terrain.argtypes = [
            C.POINTER(_TerrainParams),
            C.POINTER(_TerrainCamera),
            C.c_int,
            C.POINTER(C.c_float),
            C.POINTER(C.c_float),
            C.POINTER(C.c_float),
        ]
        lib.c3d_render_terrain.restype = C.c_int
        _scenes_lib = lib
    return _scenes_lib


def look_at(center, target, up=(0.0, 0.0, 1.0)):
    """World->camera rotation of a camera at `center` looking at `target` (x right, y down, z forward)."""
    c = np.asarray(center, float)
    zc = _unit(np.asarray(target, float) - c)
    xc = np.cross(zc, up)
    if np.linalg.norm(xc) < 1e-9:
        xc = np.cross(zc, [0.0, 1.0, 0.0])
    xc = _unit(xc)
    yc = np.cross(zc, xc)
    r = np.stack([xc, yc, zc])
    return r, -r @ c


def quantize(img: np.ndarray, noise_sigma: float = 0.0, seed: int = 0) -> np.ndarray:
    """f32 [0,1] image -> u8 via round(clamp(v, 0, 1) * 255) after optional Gaussian noise
    (the boundary image format is 8-bit, kernel.rs:87-88)."""
    v = img.astype(np.float64)
    if noise_sigma > 0:
        v = v + np.random.default_rng(seed).normal(0.0, noise_sigma, v.shape)
    return np.floor(np.clip(v, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def make_terrain_scene(
    cameras: Sequence[tuple],
    width: int,
    height: int,
    focal: float = 500.0,
    seed: int = 1,
    noise_sigma: float = 0.0,
    noise_seed: int = 2,
    extent: float = 200.0,
    relief: float = 15.0,
    names: Optional[Sequence[str]] = None,
    device: bool = False,
) -> DenseScene:
    """Renders seeded fBm terrain (extent x extent m, +-relief m) with random albedo from pinhole
    cameras given as (center, target) pairs; images quantised to u8 after optional noise. With
    `device`, the same scene model is rendered on the GPU (render.py)."""
    p = _TerrainParams(seed=seed, extent=extent, relief=relief, base_wavelength=60.0, octaves=6, albedo_scale=1.0, sky=0.7)
    cx, cy = (width - 1.0) / 2.0, (height - 1.0) / 2.0
    cams = (_TerrainCamera * len(cameras))()
    rts = []
    for i, (center, target) in enumerate(cameras):
        r, t = look_at(center, target)
        rts.append((r, t))
        cams[i].k[:] = [focal, focal, cx, cy]
        cams[i].r[:] = list(r.reshape(9))
        cams[i].t[:] = list(t)
        cams[i].width = width
        cams[i].height = height
    n = len(cameras)
    img = np.zeros((n, height, width), np.float32)
    dep = np.zeros((n, height, width), np.float32)
    nor = np.zeros((n, height, width, 3), np.float32)
    fp = C.POINTER(C.c_float)
    if device:
        from . import render

        render.render_terrain_device(p, cams, n, img, dep, nor)
    else:
        _terrain_lib().c3d_render_terrain(C.byref(p), cams, n, img.ctypes.data_as(fp), dep.ctypes.data_as(fp), nor.ctypes.data_as(fp))
    views = []
    for i, (r, t) in enumerate(rts):
        nm = names[i] if names else f"terrain_{seed}_{i}.png"
        gray = quantize(img[i], noise_sigma, noise_seed * 1000003 + i)
        views.append(DenseView(nm, width, height, np.array([focal, focal, cx, cy]), r, t, gray))
    return DenseScene(views, [], [dep[i] for i in range(n)], [nor[i] for i in range(n)])


def oblique_rig(num: int, baseline: float = 6.0, altitude: float = 30.0, ground_dist: float = 30.0, yaw_deg: float = 0.0):
    """Cameras `altitude` m up looking 45 deg down (ground_dist = altitude) at a common ground point.
    num == 3: a row across the viewing direction; otherwise a reference camera with num-1 cameras on a
    circle of radius `baseline` in the plane normal to its viewing direction."""
    yaw = math.radians(yaw_deg)
    fwd = np.array([math.sin(yaw), math.cos(yaw), 0.0])
    right = np.array([math.cos(yaw), -math.sin(yaw), 0.0])
    ref = -fwd * ground_dist + np.array([0.0, 0.0, altitude])
    target = np.zeros(3)
    vdir = _unit(target - ref)
    up = _unit(np.cross(right, vdir))
    cams = []
    if num == 3:
        for s in (0.0, -1.0, 1.0):
            cams.append((ref + s * baseline * right, target))
    else:
        cams.append((ref, target))
        for k in range(num - 1):
            a = 2 * math.pi * k / (num - 1)
            cams.append((ref + baseline * (math.cos(a) * right + math.sin(a) * up), target))
    return cams


# ---------------------------------------------------------------- kernel input


def kernel_input(
    scene: DenseScene,
    sources: Sequence[Sequence[int]],
    ranges: Sequence[tuple],
    pm: abi.PmParams,
    filt: abi.FilterParams,
    seed: int = abi.DEFAULT_SEED,
    levels: Optional[List[List[abi.LevelData]]] = None,
) -> abi.KernelInput:
    """KernelInput of a scene (densify.rs:295-312): key = FNV-1a(name), pairs from pair_geometry,
    single level unless `levels` (per view, coarsest first) is given."""
    vd = []
    for v, view in enumerate(scene.views):
        lv = levels[v] if levels is not None else [abi.LevelData(view.gray, [float(q) for q in view.k])]
        dmin, dmax = ranges[v] if ranges[v] is not None else (1.0, 2.0)
        vd.append(
            abi.ViewData(
                key=hash_bytes(view.name.encode()),
                levels=lv,
                sources=list(sources[v]),
                pairs=[pair_geometry(scene, v, s) for s in sources[v]],
                depth_min=float(np.float32(dmin)),
                depth_max=float(np.float32(dmax)),
                k=[float(q) for q in view.k],
            )
        )
    pm = abi.PmParams(**{**pm.__dict__, "sigma_spatial": pm.sigma_s()})
    return abi.KernelInput(vd, pm, filt, seed)


# ---------------------------------------------------------------- formation flights (batched workloads)


def formation_rig(positions: int, cams_per_position: int = 3, spacing: float = 8.0, altitude: float = 30.0,
                  ground_dist: float = 30.0, yaw_spread_deg: float = 30.0, cols: Optional[int] = None,
                  heading_deg: float = 0.0, offset: tuple = (0.0, 0.0)):
    """A formation flight over the terrain: `positions` stations on a serpentine grid (`cols` per
    row, `spacing` m apart) at `altitude` m, each carrying `cams_per_position` cameras that look
    45 deg down (ground_dist = altitude), fanned by +-yaw_spread_deg about the flight heading.
    The whole formation is turned by `heading_deg` about the vertical through the origin and then
    moved by `offset` (x, y) metres. Returns (center, target) pairs, station-major."""
    cols = cols or max(1, int(round(math.sqrt(positions))))
    rows = (positions + cols - 1) // cols
    cams = []
    for p in range(positions):
        r, c = divmod(p, cols)
        if r % 2:
            c = cols - 1 - c
        x = (c - (cols - 1) / 2.0) * spacing
        y = (r - (rows - 1) / 2.0) * spacing - ground_dist
        center = np.array([x, y, altitude])
        for k in range(cams_per_position):
            off = 0.0 if cams_per_position == 1 else (k / (cams_per_position - 1) - 0.5) * 2.0 * yaw_spread_deg
            yaw = math.radians(off)
            fwd = np.array([math.sin(yaw), math.cos(yaw), 0.0])
            cams.append((center, center + fwd * ground_dist - np.array([0.0, 0.0, altitude])))
    if heading_deg != 0.0 or offset[0] != 0.0 or offset[1] != 0.0:
        h = math.radians(heading_deg)
        rot = np.array([[math.cos(h), -math.sin(h), 0.0], [math.sin(h), math.cos(h), 0.0], [0.0, 0.0, 1.0]])
        shift = np.array([offset[0], offset[1], 0.0])
        cams = [(rot @ c + shift, rot @ t + shift) for c, t in cams]
    return cams


def make_mixed_scene(cameras: Sequence[tuple], sizes: Sequence[tuple], focal_1080: float = 500.0, seed: int = 1,
                     noise_sigma: float = 0.0, noise_seed: int = 2, device: bool = False) -> DenseScene:
    """Terrain views of uneven resolution: camera i is rendered at sizes[i] = (w, h) with focal
    focal_1080 * w / 1920 (the same field of view at every resolution); `device` renders on the
    GPU."""
    views = [None] * len(cameras)
    gd, gn = [None] * len(cameras), [None] * len(cameras)
    for size in sorted(set(tuple(s) for s in sizes)):
        idx = [i for i, s in enumerate(sizes) if tuple(s) == size]
        w, h = size
        sub = make_terrain_scene([cameras[i] for i in idx], w, h, focal=focal_1080 * w / 1920.0, seed=seed,
                                 noise_sigma=noise_sigma, noise_seed=noise_seed,
                                 names=[f"terrain_{seed}_{i}.png" for i in idx], device=device)
        for j, i in enumerate(idx):
            views[i], gd[i], gn[i] = sub.views[j], sub.gt_depth[j], sub.gt_normal[j]
    return DenseScene(views, [], gd, gn)


def nearest_sources(scene: DenseScene, counts: Sequence[int], min_baseline: float = 0.5) -> List[List[int]]:
    """Per view, the `counts[v]` other views whose ground footprint centres (the viewing ray's hit
    on the z = 0 plane) are closest, ties by index, among the views at least `min_baseline` m away
    (cameras of one station share a centre and cannot triangulate): overlapping views of a
    formation flight."""
    foot = []
    for v in scene.views:
        c = v.center()
        d = v.r[2]  # viewing direction in world coordinates
        t = -c[2] / d[2] if abs(d[2]) > 1e-9 else 0.0
        foot.append(c + t * d)
    out = []
    for i in range(len(scene.views)):
        ci = scene.views[i].center()
        order = sorted((float(np.linalg.norm(foot[i] - foot[j])), j) for j in range(len(scene.views))
                       if j != i and np.linalg.norm(scene.views[j].center() - ci) >= min_baseline)
        out.append([j for _, j in order[: counts[i]]])
    return out
