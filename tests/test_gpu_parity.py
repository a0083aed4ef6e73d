"""GPU parity of the gfx950 engine against the CPU oracle, through the C ABI.

Workload: 3 oblique views of 640 x 360 — seeded fBm terrain (200 x 200 m, +-15 m relief,
random albedo) imaged by 3 pinhole cameras 30 m up at 45 deg oblique, f = 500 px, 640 x 360 u8,
each view using the other two as sources, window radius 5 step 2, 5 iterations, depth 20-80 m.

Bars:
  * per-hypothesis NCC cost within 1e-4 absolute of the FP64 evaluation;
  * the engine equals the FP32 oracle bit for bit for costs, runs, evaluation, read-out,
    upsampling, the median filter and the whole device-resident schedule;
  * against the oracle computing the per-sample projective divide with IEEE division (the
    reference's literal formula; the engine uses a reproducible seeded reciprocal) depth agrees
    within 1 % relative on >= 99 % of valid pixels and normals within 2 deg on >= 99 %
    (the accuracy gate of the depth-map metric).
"""
import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes

pytestmark = pytest.mark.gpu

SRC3 = [[1, 2], [0, 2], [0, 1]]


def _c1(width=640, height=360, seed=1):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 640.0, seed=seed)
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, SRC3, [(20.0, 80.0)] * 3, pmp, fl)
    return sc, inp


@pytest.fixture(scope="module")
def c1():
    return _c1()


@pytest.fixture(scope="module")
def small():
    return _c1(192, 108, seed=4)


def _oracle(inp, rcp_mode):
    import oracle

    return oracle.Oracle(inp, threads=16, rcp_mode=rcp_mode)


def _queries(sc, v, n, seed):
    rng = np.random.default_rng(seed)
    h, w = sc.views[v].gray.shape
    xs = rng.integers(0, w, n)
    ys = rng.integers(0, h, n)
    planes = np.zeros((n, 4), np.float32)
    for i, (x, y) in enumerate(zip(xs, ys)):
        if i % 3 == 0:  # random hypothesis
            nrm = rng.normal(size=3)
            nrm[2] = -abs(nrm[2]) - 0.2
            d = rng.uniform(20, 80)
        else:  # near the true surface
            nrm = scenes.np.asarray(sc.gt_normal[v][y, x], float) + rng.normal(0, 0.05, 3)
            d = float(sc.gt_depth[v][y, x]) * rng.uniform(0.98, 1.02)
        nrm = nrm / np.linalg.norm(nrm)
        planes[i] = [*nrm, d]
    return np.stack([xs, ys], 1).astype(np.int32), planes


def test_backend_opens_gfx950():
    assert pm.is_available()
    be = pm.HipPatchMatch()
    assert be.name().startswith("hip(")


def test_pair_costs_match_fp64(c1):
    sc, inp = c1
    be = pm.HipPatchMatch()
    o = _oracle(inp, 0)
    with be.begin(inp) as s:
        for v in range(3):
            xy, planes = _queries(sc, v, 3000, seed=10 + v)
            for slot in range(2):
                got = s.pair_costs(0, v, slot, xy, planes)
                ref = np.array([o.ncc_cost_f64(0, v, slot, int(x), int(y), p) for (x, y), p in zip(xy, planes.astype(float))])
                both = (ref < 2.0) & (got < 2.0)
                assert both.sum() > 1500
                err = np.abs(got[both] - ref[both])
                print(f"view {v} slot {slot}: {both.sum()} costs, max |gpu - fp64| = {err.max():.2e}")
                assert err.max() < 1e-4
                # decisions to return 2 (outside, behind, flat) agree except at exact thresholds
                assert ((ref == 2.0) != (got == 2.0)).sum() <= 2


def test_sample_reciprocal_matches_the_oracle_over_whole_binades():
    """The window loops' per-sample reciprocal on the device (c3d_pm_sample_reciprocal) equals the
    oracle's restatement bit for bit on every float of [1, 2) and on 2^22 values over exponents
    -40..40, and stays within 1 ulp of IEEE division (the reference's `1.0f / zz`)."""
    import oracle

    be = pm.HipPatchMatch()
    z = np.arange(0x3F800000, 0x40000000, dtype=np.uint32).view(np.float32)
    rng = np.random.default_rng(5)
    z = np.concatenate([z, np.exp2(rng.uniform(-40, 40, 1 << 22)).astype(np.float32)])
    got = be.sample_reciprocal(z)
    assert np.array_equal(got.view(np.uint32), oracle.rcp_sample(z).view(np.uint32))
    ulp = np.abs(got.view(np.int32).astype(np.int64) - (np.float32(1.0) / z).view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    be.close()


def test_pair_costs_bit_exact(c1):
    sc, inp = c1
    be = pm.HipPatchMatch()
    o = _oracle(inp, 1)
    with be.begin(inp) as s:
        xy, planes = _queries(sc, 0, 2000, seed=3)
        got = s.pair_costs(0, 0, 1, xy, planes)
        ref = np.array([o.ncc_cost_f32(0, 0, 1, int(x), int(y), p) for (x, y), p in zip(xy, planes)], np.float32)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def _states(inp, views, level=0):
    out = []
    for v in views:
        h, w = inp.level_shape(v, level)
        out.append(abi.ViewState(w, h))
    return out


def _agreement(d_gpu, d_ref, n_gpu, n_ref):
    valid = (d_gpu > 0) & (d_ref > 0)
    rel = np.abs(d_gpu[valid] - d_ref[valid]) / d_ref[valid]
    cosang = np.clip((n_gpu[valid] * n_ref[valid]).sum(-1), -1, 1)
    ang = np.degrees(np.arccos(cosang))
    return valid.mean(), (rel <= 0.01).mean(), (ang <= 2.0).mean()


def test_run_bit_exact(c1):
    """Random-init photometric run, 5 iterations (3 views of 640 x 360): identical to the oracle."""
    sc, inp = c1
    be = pm.HipPatchMatch()
    o = _oracle(inp, 1)
    p = abi.run_params(0, False, True, 5, 1, False)
    g = _states(inp, [0, 1, 2])
    r = _states(inp, [0, 1, 2])
    with be.begin(inp) as s:
        s.run([0, 1, 2], g, p)
    o.run([0, 1, 2], r, p)
    for a, b in zip(g, r):
        assert np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32))
        assert np.array_equal(a.normal.view(np.uint32), b.normal.view(np.uint32))
        assert np.array_equal(a.cost.view(np.uint32), b.cost.view(np.uint32))


def test_run_within_tolerance_of_ieee_divide(c1):
    sc, inp = c1
    be = pm.HipPatchMatch()
    o = _oracle(inp, 0)
    p = abi.run_params(0, False, True, 5, 1, False)
    g = _states(inp, [0, 1, 2])
    r = _states(inp, [0, 1, 2])
    with be.begin(inp) as s:
        s.run([0, 1, 2], g, p)
    o.run([0, 1, 2], r, p)
    for v, (a, b) in enumerate(zip(g, r)):
        valid, dfrac, nfrac = _agreement(a.depth, b.depth, a.normal, b.normal)
        print(f"view {v}: valid {valid:.4f} depth<=1% {dfrac:.5f} normal<=2deg {nfrac:.5f}")
        assert dfrac >= 0.99 and nfrac >= 0.99


def test_evaluate_filter_bit_exact(small):
    sc, inp = small
    be = pm.HipPatchMatch()
    o = _oracle(inp, 1)
    views = [0, 1, 2]
    g = _states(inp, views)
    p = abi.run_params(0, False, True, 3, 1, False)
    o.run(views, g, p)
    snap = abi.DepthSnapshot(7, 0, [s.depth for s in g])
    with be.begin(inp) as s:
        ce = s.evaluate(0, views, g, False)
        cg = s.evaluate(0, views, g, True, snap)
        cnt = s.filter(views, g, snap)
        # geometric run from the same states
        gg = [x.copy() for x in g]
        s.run(views, gg, abi.run_params(0, True, False, 2, 2, False), snap)
    oe = o.evaluate(0, views, g, False)
    og = o.evaluate(0, views, g, True, snap)
    oc = o.filter(views, g, snap)
    og2 = [x.copy() for x in g]
    o.run(views, og2, abi.run_params(0, True, False, 2, 2, False), snap)
    for i in range(3):
        assert np.array_equal(ce[i].view(np.uint32), oe[i].view(np.uint32))
        assert np.array_equal(cg[i].view(np.uint32), og[i].view(np.uint32))
        assert np.array_equal(cnt[i], oc[i])
        assert np.array_equal(gg[i].depth.view(np.uint32), og2[i].depth.view(np.uint32))
        assert np.array_equal(gg[i].normal.view(np.uint32), og2[i].normal.view(np.uint32))
    assert sum(int(c.sum()) for c in cnt) > 0


def test_densify_within_tolerance_of_ieee_divide(small):
    """Device-resident single-level schedule (photometric 5, one geometric round of 2, read-out)."""
    sc, inp = small
    be = pm.HipPatchMatch()
    o = _oracle(inp, 0)
    with be.begin(inp) as s:
        go, gc, graw, tm = s.densify([0, 1, 2], abi.Schedule(photometric_iters=5, geometric_rounds=1, geometric_iters=2))
    ro, rc, rraw = o.densify([0, 1, 2], abi.Schedule(photometric_iters=5, geometric_rounds=1, geometric_iters=2))
    assert tm.total_ms > 0 and tm.half_launches == 2 * (5 + 2)
    for v in range(3):
        valid, dfrac, nfrac = _agreement(graw[v], rraw[v], go[v].normal, ro[v].normal)
        print(f"view {v}: raw valid {valid:.4f} depth<=1% {dfrac:.5f}")
        assert dfrac >= 0.99
        # the read-out counts agree on nearly all pixels
        assert (gc[v] == rc[v]).mean() > 0.98


def test_densify_bit_exact(small):
    """c3d_pm_densify (states resident on the device) equals the oracle's schedule bit for bit."""
    sc, inp = small
    be = pm.HipPatchMatch()
    o = _oracle(inp, 1)
    with be.begin(inp) as s:
        go, gc, graw, tm = s.densify([0, 1, 2], abi.Schedule(photometric_iters=3, geometric_rounds=1, geometric_iters=2))
    ro, rc, rraw = o.densify([0, 1, 2], abi.Schedule(photometric_iters=3, geometric_rounds=1, geometric_iters=2))
    for v in range(3):
        assert np.array_equal(graw[v].view(np.uint32), rraw[v].view(np.uint32))
        assert np.array_equal(gc[v], rc[v])
        assert np.array_equal(go[v].depth.view(np.uint32), ro[v].depth.view(np.uint32))


def _two_level(seed=5):
    import oracle

    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, 160, 96, focal=125.0, seed=seed)
    levels = []
    L = oracle.lib()
    import ctypes as C

    for v in sc.views:
        low = np.zeros((48, 80), np.uint8)
        outs = (C.POINTER(C.c_uint8) * 1)(low.ctypes.data_as(C.POINTER(C.c_uint8)))
        ko = (C.c_float * 8)()
        L.pmo_build_pyramid(v.gray.ctypes.data_as(C.POINTER(C.c_uint8)), 160, 96, (C.c_double * 4)(*v.k), 2, outs, ko)
        levels.append([abi.LevelData(low, list(ko[4:8])), abi.LevelData(v.gray, list(ko[0:4]))])
    pmp, fl = abi.fast_profile()
    return sc, scenes.kernel_input(sc, SRC3, [(20.0, 80.0)] * 3, pmp, fl, levels=levels)


def test_upsample_median_bit_exact():
    sc, inp = _two_level()
    be = pm.HipPatchMatch()
    o = _oracle(inp, 1)
    views = [0, 1, 2]
    low = _states(inp, views, 0)
    o.run(views, low, abi.run_params(0, False, True, 3, 1, False))
    with be.begin(inp) as s:
        up = s.upsample(1, views, low, 1.0, 0.04)
        med = s.median_filter(1, views, up)
    oup = o.upsample(1, views, low, 1.0, 0.04)
    omed = o.median_filter(1, views, oup)
    for i in range(3):
        assert np.array_equal(up[i].depth.view(np.uint32), oup[i].depth.view(np.uint32))
        assert np.array_equal(up[i].normal.view(np.uint32), oup[i].normal.view(np.uint32))
        assert np.array_equal(med[i].depth.view(np.uint32), omed[i].depth.view(np.uint32))
    assert (up[0].depth > 0).mean() > 0.9


def test_errors_map_to_invalid_argument(small):
    sc, inp = small
    be = pm.HipPatchMatch()
    bad = abi.KernelInput(inp.views, abi.PmParams(window_radius=21), inp.filter)
    with pytest.raises(pm.PmError) as e:
        be.begin(bad)
    assert e.value.status == abi.C3D_INVALID_ARGUMENT
    with be.begin(inp) as s:
        wrong = [abi.ViewState(10, 10)]
        with pytest.raises(pm.PmError) as e:
            s.run([0], wrong, abi.run_params(0, False, True, 1, 1, False))
        assert e.value.status == abi.C3D_INVALID_ARGUMENT
        with pytest.raises(pm.PmError) as e:
            s.run([0], _states(inp, [0]), abi.run_params(0, True, False, 1, 1, False), None)
        assert e.value.status == abi.C3D_INVALID_ARGUMENT


@pytest.mark.parametrize("sizes", [[(1, 1), (3, 2), (2, 3)], [(13, 7), (130, 5), (7, 40)], [(129, 65), (64, 1), (257, 3)]])
def test_tiny_and_ragged_views_bit_exact(sizes):
    """Views smaller than the window, one pixel wide or high, and widths off every tile size, in
    one batch: a random-init run, a geometric iteration and the read-out equal the oracle's."""
    import oracle

    cams = scenes.formation_rig(1, 3)
    sc = scenes.make_mixed_scene(cams, sizes, seed=9)
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, SRC3, [(15.0, 120.0)] * 3, pmp, fl)
    views = [0, 1, 2]
    be = pm.HipPatchMatch()
    g = [abi.ViewState(w, h) for (w, h) in sizes]
    with be.begin(inp) as s:
        s.run(views, g, abi.run_params(0, False, True, 2, 1, False))
        snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in g])
        s.run(views, g, abi.run_params(0, True, False, 1, 2, False), snap)
        gc = s.filter(views, g, snap)
    be.close()
    o = oracle.Oracle(inp, threads=4)
    r = [abi.ViewState(w, h) for (w, h) in sizes]
    o.run(views, r, abi.run_params(0, False, True, 2, 1, False))
    rsnap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in r])
    o.run(views, r, abi.run_params(0, True, False, 1, 2, False), rsnap)
    rc = o.filter(views, r, rsnap)
    for v in views:
        for a, b in ((g[v].depth, r[v].depth), (g[v].normal, r[v].normal), (g[v].cost, r[v].cost)):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
        assert np.array_equal(gc[v], rc[v])


@pytest.mark.parametrize("cams,size", [(12, (96, 64)), (33, (40, 24))])
def test_many_sources_bit_exact_every_tile_shape(cams, size):
    """Every view using all others as sources (11, and the maximum of 32): the refinement's
    source order and early exit, and the LDS tables sized by the source count, under each
    half-step workgroup shape; states, costs and the read-out equal the oracle's."""
    import oracle

    rig = scenes.formation_rig(cams // 3, 3)
    sc = scenes.make_mixed_scene(rig, [size] * cams, seed=11)
    srcs = [[j for j in range(cams) if j != i] for i in range(cams)]
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, srcs, [(15.0, 120.0)] * cams, pmp, fl)
    assert max(len(s) for s in srcs) <= 32  # C3D_MAX_SOURCES
    views = list(range(cams))

    def run(engine):
        st = [abi.ViewState(*size) for _ in views]
        engine.run(views, st, abi.run_params(0, False, True, 2, 1, False))
        snap = abi.DepthSnapshot(cams, 0, [x.depth.copy() for x in st])
        engine.run(views, st, abi.run_params(0, True, False, 1, 2, False), snap)
        return st, engine.filter(views, st, snap)

    o = oracle.Oracle(inp, threads=16)
    ref, rc = run(o)
    for rows, tpp in abi.TILE_SHAPES:
        opts = abi.c3d_pm_options()
        opts.tile_rows, opts.threads_per_pixel = rows, tpp
        be = pm.HipPatchMatch(0, opts)
        with be.begin(inp) as s:
            got, gc = run(s)
        be.close()
        for a, b in zip(got, ref):
            for x, y in ((a.depth, b.depth), (a.normal, b.normal), (a.cost, b.cost)):
                assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32)), (rows, tpp)
        for a, b in zip(gc, rc):
            assert np.array_equal(a, b), (rows, tpp)


def test_pair_costs_bit_exact_at_the_maximum_image_size():
    """Views of 4093 x 4093, the largest the source tables address exactly ((w + 3)(h + 3) = 2^24):
    per-hypothesis costs at pixels near every corner and in the middle, random planes, equal the
    oracle's bit for bit (uniform-noise images; the geometry of the 3-view oblique rig)."""
    import oracle

    side = 4093
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, 64, 48, focal=50.0, seed=3)
    rng = np.random.default_rng(17)
    f = 50.0 * side / 64.0
    levels = [[abi.LevelData(rng.integers(0, 256, (side, side), dtype=np.uint8), [f, f, (side - 1) / 2.0, (side - 1) / 2.0])]
              for _ in sc.views]
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, SRC3, [(20.0, 80.0)] * 3, pmp, fl, levels=levels)
    n = 1500
    edge = np.concatenate([rng.integers(0, 64, n // 3), rng.integers(side - 64, side, n // 3),
                           rng.integers(0, side, n - 2 * (n // 3))])
    xy = np.stack([edge, rng.permutation(edge)], 1).astype(np.int32)
    nrm = rng.normal(size=(n, 3))
    nrm[:, 2] = -np.abs(nrm[:, 2]) - 0.5
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    planes = np.concatenate([nrm, rng.uniform(20.0, 80.0, (n, 1))], 1).astype(np.float32)
    be = pm.HipPatchMatch()
    with be.begin(inp) as s:
        got = s.pair_costs(0, 0, 1, xy, planes)
    be.close()
    o = oracle.Oracle(inp, threads=16)
    ref = np.array([o.ncc_cost_f32(0, 0, 1, int(x), int(y), p) for (x, y), p in zip(xy, planes)], np.float32)
    assert (ref < 2.0).sum() > n // 10  # enough windows inside the source to test its far addresses
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("random_init", [0, 2])
def test_caller_states_with_nan_zero_and_negative_depths(small, random_init):
    """Caller-provided states holding NaN, zero and negative depths and NaN normals, run without
    random initialisation (such planes cost 2 and are replaced by propagation) and with
    random_init = 2 (only pixels without a positive depth are re-initialised): 2 iterations and
    the read-out equal the oracle's bit for bit."""
    import oracle

    sc, inp = small
    views = [0, 1, 2]

    def states():
        out = []
        for v in views:
            h, w = inp.level_shape(v, 0)
            st = abi.ViewState(w, h)
            st.depth[...] = np.asarray(sc.gt_depth[v], np.float32) * np.float32(1.03)
            st.normal[...] = np.asarray(sc.gt_normal[v], np.float32)
            r = np.random.default_rng(100 + v).random((h, w))
            st.depth[r < 0.05] = np.nan
            st.depth[(r >= 0.05) & (r < 0.10)] = 0.0
            st.depth[(r >= 0.10) & (r < 0.15)] = -5.0
            st.normal[(r >= 0.15) & (r < 0.18)] = np.nan
            out.append(st)
        return out

    p = abi.run_params(0, False, random_init, 2, 3, False)
    be = pm.HipPatchMatch()
    g = states()
    with be.begin(inp) as s:
        s.run(views, g, p)
        snap = abi.DepthSnapshot(7, 0, [x.depth.copy() for x in g])
        gc = s.filter(views, g, snap)
    be.close()
    o = oracle.Oracle(inp, threads=16)
    r = states()
    o.run(views, r, p)
    rc = o.filter(views, r, abi.DepthSnapshot(7, 0, [x.depth.copy() for x in r]))
    for v in views:
        for a, b in ((g[v].depth, r[v].depth), (g[v].normal, r[v].normal), (g[v].cost, r[v].cost)):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
        assert np.array_equal(gc[v], rc[v])


def test_source_less_view_and_repeated_source():
    """A view without sources (its planes cost 2 and nothing is adopted) and a view listing one
    source twice, in one batch with a normal view: a random-init run, a geometric iteration and
    the read-out equal the oracle's bit for bit."""
    import oracle

    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, 160, 96, focal=500.0 * 160 / 640.0, seed=6)
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, [[], [0, 0, 2], [0, 1]], [(20.0, 80.0)] * 3, pmp, fl)
    views = [0, 1, 2]

    def run(engine):
        st = [abi.ViewState(160, 96) for _ in views]
        engine.run(views, st, abi.run_params(0, False, True, 2, 1, False))
        snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in st])
        engine.run(views, st, abi.run_params(0, True, False, 1, 2, False), snap)
        return st, engine.filter(views, st, snap)

    be = pm.HipPatchMatch()
    with be.begin(inp) as s:
        got, gc = run(s)
    be.close()
    ref, rc = run(oracle.Oracle(inp, threads=16))
    for a, b in zip(got, ref):
        for x, y in ((a.depth, b.depth), (a.normal, b.normal), (a.cost, b.cost)):
            assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    for a, b in zip(gc, rc):
        assert np.array_equal(a, b)
    assert np.all(gc[0] == 0)  # nothing to agree with
