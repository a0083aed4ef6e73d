"""Error behaviour of the post-processing, scene-preparation and densify entry points on a device:
arguments the reference rejects (or that would index out of bounds) return INVALID_ARGUMENT with a
message, and nothing is launched. Mirrors the reference's InvalidArgument paths
(patchmatch.rs:292-294,366-371,640-642,754; densify.rs:221-223)."""
import ctypes as C

import numpy as np
import pytest

from cumulus3d_amd import abi, neighbors, pm, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = pm.HipPatchMatch()
    yield b
    b.close()


def _status(fn):
    with pytest.raises(pm.PmError) as e:
        fn()
    return e.value.status, e.value.message


def test_post_arguments(be):
    h, w = 8, 8
    d = np.ones((h, w), np.float32)
    n = np.zeros((h, w, 3), np.float32)
    c = np.zeros((h, w), np.float32)
    st, msg = _status(lambda: be.fill_holes(d, n, c, c, np.zeros((h, w), np.uint8), [10, 10, 4, 4],
                                            abi.PostParams(fill_holes=True, fill_max_frac=2.0)))
    assert st == abi.C3D_INVALID_ARGUMENT and "fill_max_frac" in msg
    st, _ = _status(lambda: be.remove_speckles(d, n, c, abi.PostParams(speckle_min_area_frac=-1.0)))
    assert st == abi.C3D_INVALID_ARGUMENT
    L = be.L
    p = abi.PostParams().c()
    fp = C.POINTER(C.c_float)
    rc = L.c3d_fill_holes(be.h, d.ctypes.data_as(fp), n.ctypes.data_as(fp), c.ctypes.data_as(fp), c.ctypes.data_as(fp), None,
                          (C.c_double * 4)(10, 10, 4, 4), w, h, C.byref(p), None)
    assert rc == abi.C3D_INVALID_ARGUMENT
    rc = L.c3d_remove_speckles(be.h, d.ctypes.data_as(fp), n.ctypes.data_as(fp), c.ctypes.data_as(fp), 0, h, C.byref(p), None)
    assert rc == abi.C3D_INVALID_ARGUMENT


def test_scene_arguments(be):
    r = np.diag([1.0, -1.0, -1.0])
    v = scenes.DenseView("v", 4, 4, np.array([10.0, 10.0, 1.5, 1.5]), r, np.array([0.0, 0.0, 10.0]), np.zeros((4, 4), np.uint8))
    sc = scenes.DenseScene([v, v], [scenes.ScenePoint(np.zeros(3), [0, 2])], None, None)  # view 2 does not exist
    st, msg = _status(lambda: neighbors.pair_stats(be, sc))
    assert st == abi.C3D_INVALID_ARGUMENT and "out of range" in msg
    st, _ = _status(lambda: neighbors.depth_ranges(be, sc))
    assert st == abi.C3D_INVALID_ARGUMENT
    s = neighbors.SparseScene(scenes.DenseScene([v, v], [scenes.ScenePoint(np.zeros(3), [0, 1])] * 2, None, None))
    s.off[1] = 5  # offsets must not decrease (5 then 4)
    st, msg = _status(lambda: neighbors.pair_stats(be, s))
    assert st == abi.C3D_INVALID_ARGUMENT and "track_offset" in msg
    big = scenes.DenseScene([v] * 8193, [], None, None)
    st, msg = _status(lambda: neighbors.pair_stats(be, big))
    assert st == abi.C3D_INVALID_ARGUMENT and "8192" in msg
    st, _ = _status(lambda: neighbors.select(np.zeros((2, 2), np.uint32), np.zeros((2, 2)), 5, 3, 0.0))
    assert st == abi.C3D_INVALID_ARGUMENT


def test_densify_post_needs_intrinsics(be):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, 64, 48, focal=50.0, seed=3)
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, [[1, 2], [0, 2], [0, 1]], [(20.0, 80.0)] * 3, pmp, fl)
    sch = abi.Schedule(post=abi.PostParams.fast()).c(None)  # fill_holes without post_k
    with be.begin(inp) as s:
        rc = s.L.c3d_pm_densify(s.h, abi.views_array([0, 1, 2]), 3, C.byref(sch), None, None, None, None)
        assert rc == abi.C3D_INVALID_ARGUMENT
        assert "post_k" in s.L.c3d_pm_last_error().decode()
        # the session stays usable
        outs, _, _, tm = s.densify([0, 1, 2], abi.Schedule(photometric_iters=1, geometric_rounds=0, post=abi.PostParams.fast()))
        assert all(o.depth.shape == (48, 64) for o in outs)


def test_empty_view_lists_are_noops(be):
    """No views: run, evaluate, filter, upsample, median and densify do nothing and succeed (the
    reference's batches and schedule loop over no views, patchmatch.rs:606-625, densify.rs:290-425);
    the autotuner, which has nothing to time, refuses; the session stays usable."""
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, 128, 96, focal=100.0, seed=3)
    pmp, fl = abi.fast_profile()
    levels = [be.build_pyramid(v.gray, v.k, 2) for v in sc.views]
    inp = scenes.kernel_input(sc, [[1, 2], [0, 2], [0, 1]], [(20.0, 80.0)] * 3, pmp, fl, levels=levels)
    assert inp.num_levels == 2
    snap = abi.DepthSnapshot(1, 1, [None, None, None])
    with be.begin(inp) as s:
        s.run([], [], abi.run_params(0, False, True, 1, 1, False))
        assert s.evaluate(1, [], []) == []
        assert s.filter([], [], snap) == []
        assert s.upsample(1, [], [], 1.0, 10.0) == []
        assert s.median_filter(1, [], []) == []
        outs, counts, raws, tm = s.densify([], abi.Schedule(photometric_iters=1, geometric_rounds=0))
        assert outs == [] and counts == [] and raws == [] and tm.total_ms == 0.0
        st, msg = _status(lambda: s.autotune([]))
        assert st == abi.C3D_INVALID_ARGUMENT and "no views" in msg
        outs, _, _, _ = s.densify([0, 1, 2], abi.Schedule(photometric_iters=1, geometric_rounds=0))
        assert all(o.depth.shape == (96, 128) for o in outs)


def test_image_size_limit(be):
    """Source-table byte offsets are formed exactly in FP32, so views are limited to
    (w + 3)(h + 3) <= 2^24 pixels: 4093 x 4093 opens, 4096 x 4096 is refused at begin."""
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, 64, 48, focal=50.0, seed=3)
    pmp, fl = abi.fast_profile()
    for side, ok in ((4093, True), (4096, False)):
        big = np.zeros((side, side), np.uint8)
        levels = [[abi.LevelData(big, [50.0, 50.0, side / 2.0, side / 2.0])] for _ in sc.views]
        inp = scenes.kernel_input(sc, [[1, 2], [0, 2], [0, 1]], [(20.0, 80.0)] * 3, pmp, fl, levels=levels)
        if ok:
            be.begin(inp).close()
        else:
            st, msg = _status(lambda: be.begin(inp))
            assert st == abi.C3D_INVALID_ARGUMENT and "2^24" in msg
