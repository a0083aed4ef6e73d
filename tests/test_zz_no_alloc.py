"""c3d_fill_holes_noalloc (the allocation-free gap filling) against c3d_fill_holes through the C ABI:
bit for bit on a 1920 x 1080 depth map with holes and then on a smaller one (the fixed device
scratch keeps the first call's contents), and images above its 2^21-pixel limit are refused."""
import ctypes as C

import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes

pytestmark = pytest.mark.gpu


def _case(width, height, seed):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 1920.0, seed=seed, device=True)
    rng = np.random.default_rng(seed)
    depth = sc.gt_depth[0].copy()
    normal = sc.gt_normal[0].copy()
    yy, xx = np.mgrid[0:height, 0:width]
    for _ in range(40):  # holes of up to 30 px radius
        cx, cy, r = rng.integers(0, width), rng.integers(0, height), rng.integers(2, 30)
        hole = (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r
        depth[hole] = 0.0
        normal[hole] = 0.0
    cost = rng.uniform(0.0, 1.2, (height, width)).astype(np.float32)
    raw = rng.uniform(0.0, 1.2, (height, width)).astype(np.float32)
    return depth, normal, cost, raw, sc.views[0].gray, [float(q) for q in sc.views[0].k]


def _call(be, fn, depth, normal, cost, raw, gray, k, params):
    d, n, c = depth.copy(), normal.copy(), cost.copy()
    h, w = d.shape
    cnt = C.c_uint64()
    fp = C.POINTER(C.c_float)
    p = params.c()
    rc = getattr(be.L, fn)(be.h, d.ctypes.data_as(fp), n.ctypes.data_as(fp), c.ctypes.data_as(fp), raw.ctypes.data_as(fp),
                           np.ascontiguousarray(gray).ctypes.data_as(C.POINTER(C.c_uint8)), (C.c_double * 4)(*k), w, h, C.byref(p),
                           C.byref(cnt))
    return rc, d, n, c, int(cnt.value)


def test_noalloc_fill_holes_matches_fill_holes():
    be = pm.HipPatchMatch()
    params = abi.PostParams(fill_holes=True, fill_max_frac=0.02, fill_edge_step=0.1, fill_max_cost=0.9)
    for width, height, seed in ((1920, 1080, 3), (640, 360, 4)):
        case = _case(width, height, seed)
        rc0, d0, n0, c0, k0 = _call(be, "c3d_fill_holes", *case, params)
        rc1, d1, n1, c1, k1 = _call(be, "c3d_fill_holes_noalloc", *case, params)
        assert rc0 == abi.C3D_OK and rc1 == abi.C3D_OK
        print(f"{width}x{height}: {k0} pixels filled")
        assert k0 > 1000 and k1 == k0
        assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
        assert np.array_equal(n0.view(np.uint32), n1.view(np.uint32))
        assert np.array_equal(c0.view(np.uint32), c1.view(np.uint32))
    # above the fixed scratch: refused before any device work, the caller's arrays untouched
    d = np.zeros((1024, 2049), np.float32)
    rc, d2, _, _, _ = _call(be, "c3d_fill_holes_noalloc", d, np.zeros((1024, 2049, 3), np.float32), d, d,
                            np.zeros((1024, 2049), np.uint8), [500.0, 500.0, 1024.0, 511.5], params)
    assert rc == abi.C3D_INVALID_ARGUMENT and b"2^21" in be.L.c3d_pm_last_error()
    be.close()
