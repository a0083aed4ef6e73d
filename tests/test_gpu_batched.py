"""Batched views of uneven resolution and source count in one flattened launch per half-step, and
the whole multi-scale schedule captured into one hipGraph (c3d_pm_options.use_graph), against the
CPU oracle and the eager path, through the C ABI.

Bars: the batched mixed-resolution schedule equals the oracle bit for bit; graph replays equal the
eager schedule bit for bit, replay after replay with eager work between them (a graph keeps its
own copies of every parameter), and a new view set is captured anew.
"""
import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes

pytestmark = pytest.mark.gpu

SIZES = [(256, 144), (320, 180), (384, 216)]


def _mixed(seed=3):
    rng = np.random.default_rng(seed)
    cams = scenes.formation_rig(3, 3)
    sizes = [SIZES[k] for k in rng.integers(0, 3, len(cams))]
    sc = scenes.make_mixed_scene(cams, sizes, seed=seed)
    src = scenes.nearest_sources(sc, [int(k) for k in rng.integers(2, 5, len(cams))])
    return sc, src


def _input(be, sc, src):
    nl = pm.num_levels(min(min(v.width, v.height) for v in sc.views), 3, 64)
    levels = [be.build_pyramid(v.gray, v.k, nl) for v in sc.views]
    pmp, fl = abi.fast_profile()
    return scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(sc.views), pmp, fl, levels=levels)


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


SCHED = abi.Schedule(photometric_iters=3, photometric_iters_upper=2, geometric_rounds=1, geometric_iters=2)


def test_mixed_resolution_batch_bit_exact():
    import oracle

    sc, src = _mixed()
    be = pm.HipPatchMatch()
    inp = _input(be, sc, src)
    assert inp.num_levels == 2 and len({v.width for v in sc.views}) > 1
    views = list(range(len(sc.views)))
    with be.begin(inp) as s:
        go, gc, graw, _ = s.densify(views, SCHED)
    ro, rc, rraw = oracle.Oracle(inp, threads=16).densify(views, SCHED)
    for v in views:
        assert go[v].depth.shape == (sc.views[v].height, sc.views[v].width)
        assert _same(graw[v], rraw[v])
        assert np.array_equal(gc[v], rc[v])
        assert _same(go[v].normal, ro[v].normal)
    assert np.mean([(g.depth > 0).mean() for g in go]) > 0.3


def test_graph_replays_equal_eager():
    sc, src = _mixed(seed=4)
    views = list(range(len(sc.views)))
    eager_be = pm.HipPatchMatch()
    inp = _input(eager_be, sc, src)
    with eager_be.begin(inp) as s:
        eo, ec, eraw, etm = s.densify(views, SCHED)
        sub = [4, 1, 2]
        so, sc_, sraw, _ = s.densify(sub, SCHED)
    eager_be.close()
    opts = abi.c3d_pm_options()
    opts.use_graph = 1
    be = pm.HipPatchMatch(0, opts)
    with be.begin(inp) as s:
        for rep in range(3):
            go, gc, graw, tm = s.densify(views, SCHED)
            assert tm.half_launches == etm.half_launches and tm.total_ms > 0
            for v in views:
                assert _same(graw[v], eraw[v]) and np.array_equal(gc[v], ec[v]) and _same(go[v].normal, eo[v].normal)
            # eager work on the same session between replays
            st = [abi.ViewState(*inp.level_shape(v, 0)[::-1]) for v in views]
            s.evaluate(0, views, st, False)
        # another view set is captured anew
        go, gc, graw, _ = s.densify(sub, SCHED)
        for i in range(len(sub)):
            assert _same(graw[i], sraw[i]) and np.array_equal(gc[i], sc_[i])


def test_two_lanes_per_pixel_graph_equals_default():
    """The mixed batch's schedule captured with two lanes per pixel (tile autotuner's (4, 2)
    shape) replays to the default eager shape's results bit for bit."""
    sc, src = _mixed(seed=5)
    views = list(range(len(sc.views)))
    be = pm.HipPatchMatch()
    inp = _input(be, sc, src)
    with be.begin(inp) as s:
        eo, ec, eraw, _ = s.densify(views, SCHED)
    be.close()
    opts = abi.c3d_pm_options()
    opts.use_graph, opts.tile_rows, opts.threads_per_pixel = 1, 4, 2
    be = pm.HipPatchMatch(0, opts)
    with be.begin(inp) as s:
        for _ in range(2):
            go, gc, graw, _ = s.densify(views, SCHED)
            for v in views:
                assert _same(graw[v], eraw[v]) and np.array_equal(gc[v], ec[v]) and _same(go[v].normal, eo[v].normal)
    be.close()
