"""GPU parity of the multi-scale schedule (compute_depth_maps, crates/dense/src/densify.rs:350-454)
and of the device gray pyramid (build_pyramid, crates/dense/src/upsample.rs:78-102), through the
C ABI.

Bars:
  * the device pyramid equals the oracle's byte for byte, intrinsics included;
  * c3d_pm_densify over 3 scales (fast profile: 6 / 3 photometric iterations, detail restorer,
    2 geometric rounds of 2 per scale) equals the oracle's schedule bit for bit; so does the
    quality-profile variant with the restorer's random init, the weak-texture prior and the median
    filter switched on;
  * on the reference's own ground-truth scene (SynthConfig::default(): 3 x 3 cameras, 384 x 288,
    3 scales) the bars of crates/cuda/tests/patchmatch_gpu.rs:96-98 hold for both profiles
    (depth median relative error < 0.005, normal median < 3 deg, valid > 0.5), and a second run is
    bit-identical (patchmatch_gpu.rs:112-117).
"""
import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes

pytestmark = pytest.mark.gpu


def _oracle(inp):
    import oracle

    return oracle.Oracle(inp, threads=16)


def _device_levels(be, sc, nl):
    return [be.build_pyramid(v.gray, v.k, nl) for v in sc.views]


def _terrain(width, height, nviews=3, seed=7):
    rig = scenes.oblique_rig(nviews, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 640.0, seed=seed)
    src = [[j for j in range(nviews) if j != v] for v in range(nviews)]
    return sc, src


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_device_pyramid_bit_exact():
    import oracle

    be = pm.HipPatchMatch()
    rng = np.random.default_rng(3)
    for (h, w, levels) in [(61, 97, 3), (288, 384, 3), (64, 64, 1), (33, 130, 4)]:
        img = rng.integers(0, 256, (h, w)).astype(np.uint8)
        k = [300.0, 310.0, (w - 1) / 2.0 + 0.25, (h - 1) / 2.0 - 0.5]
        dev = be.build_pyramid(img, k, levels)
        ref = oracle.build_pyramid(img, k, levels)
        assert len(dev) == levels
        for d, (rg, rk) in zip(dev, ref):
            assert d.gray.shape == rg.shape
            assert np.array_equal(d.gray, rg)
            assert [np.float32(q) for q in d.k] == [np.float32(q) for q in rk]
    assert pm.num_levels(539, 3, 64) == 3 and pm.num_levels(240, 3, 64) == 2 and pm.num_levels(90, 3, 64) == 1


def test_multiscale_densify_bit_exact():
    """Three scales of a 320 x 256 terrain (80 x 64, 160 x 128, 320 x 256), fast profile."""
    sc, src = _terrain(320, 256)
    be = pm.HipPatchMatch()
    nl = pm.num_levels(256, 3, 64)
    assert nl == 3
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, [(20.0, 80.0)] * 3, pmp, fl, levels=_device_levels(be, sc, nl))
    sched = abi.Schedule.fast()
    with be.begin(inp) as s:
        go, gc, graw, tm = s.densify([0, 1, 2], sched)
    ro, rc, rraw = _oracle(inp).densify([0, 1, 2], sched)
    # 6 + 3 + 3 photometric and 3 x 2 x 2 geometric iterations, two half-steps each
    assert tm.half_launches == 2 * (6 + 3 + 3 + 3 * 2 * 2)
    for v in range(3):
        print(f"view {v}: valid {(go[v].depth > 0).mean():.3f} oracle {(ro[v].depth > 0).mean():.3f}")
        assert _same(graw[v], rraw[v])
        assert np.array_equal(gc[v], rc[v])
        assert _same(go[v].depth, ro[v].depth)
        assert _same(go[v].normal, ro[v].normal)
        assert _same(go[v].cost, ro[v].cost)
    assert (go[0].depth > 0).mean() > 0.5


def test_multiscale_quality_prior_median_bit_exact():
    """Two scales, restorer random init, weak-texture prior and median filter on."""
    sc, src = _terrain(192, 128, seed=11)
    be = pm.HipPatchMatch()
    pmp, fl = abi.fast_profile()
    pmp = abi.PmParams(**{**pmp.__dict__, "weak_texture_var": 0.004})
    fl = abi.FilterParams(**{**fl.__dict__, "median_filter": True})
    inp = scenes.kernel_input(sc, src, [(20.0, 80.0)] * 3, pmp, fl, levels=_device_levels(be, sc, 2))
    sched = abi.Schedule(photometric_iters=4, photometric_iters_upper=3, geometric_rounds=1, geometric_iters=2,
                         restorer_random_init=True, first_run_id=5)
    with be.begin(inp) as s:
        go, gc, graw, _ = s.densify([2, 0, 1], sched)
    ro, rc, rraw = _oracle(inp).densify([2, 0, 1], sched)
    for v in range(3):
        assert _same(graw[v], rraw[v])
        assert np.array_equal(gc[v], rc[v])
        assert _same(go[v].normal, ro[v].normal)


def _accuracy(sc, outs, views):
    """accuracy() of crates/cuda/tests/patchmatch_gpu.rs:50-76."""
    de, ne, total = [], [], 0
    for i, v in enumerate(views):
        d = outs[i].depth.reshape(-1).astype(np.float64)
        t = sc.gt_depth[v].reshape(-1).astype(np.float64)
        total += d.size
        m = (d > 0) & (t > 0)
        de.append(np.abs(d[m] - t[m]) / t[m])
        c = (outs[i].normal.reshape(-1, 3)[m].astype(np.float64) * sc.gt_normal[v].reshape(-1, 3)[m]).sum(1)
        ne.append(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))
    de, ne = np.concatenate(de), np.concatenate(ne)
    return de.size / total, float(np.median(de)), float(np.median(ne))


@pytest.fixture(scope="module")
def synth():
    from cumulus3d_amd import neighbors

    sc = scenes.make_synth_scene()
    be = pm.HipPatchMatch()
    src, ranges = neighbors.sources_and_ranges(be, sc)
    be.close()
    return sc, src, ranges


@pytest.mark.parametrize("profile", ["fast", "quality"])
def test_reference_ground_truth_bars(synth, profile):
    sc, src, ranges = synth
    be = pm.HipPatchMatch()
    h, w = sc.views[0].gray.shape
    nl = pm.num_levels(min(w, h), 3, 64)
    assert nl == 3
    pmp, fl = abi.fast_profile() if profile == "fast" else abi.quality_profile()
    inp = scenes.kernel_input(sc, src, ranges, pmp, fl, levels=_device_levels(be, sc, nl))
    sched = abi.Schedule.fast() if profile == "fast" else abi.Schedule.quality()
    views = list(range(len(sc.views)))
    with be.begin(inp) as s:
        outs, _, raws, tm = s.densify(views, sched)
        again, _, raws2, _ = s.densify(views, sched)
    valid, dmed, nmed = _accuracy(sc, outs, views)
    print(f"{profile}: valid {valid:.3f} depth median {dmed:.2e} normal median {nmed:.2f} deg, {tm.total_ms:.1f} ms")
    assert dmed < 0.005
    assert nmed < 3.0
    assert valid > 0.5
    for a, b, c, d in zip(raws, raws2, outs, again):
        assert _same(a, b) and _same(c.normal, d.normal)


def test_reference_fusion_surface(synth):
    """patchmatch_gpu.rs:119-132: the consistency fusion of the fast profile's depth maps (with its
    speckle removal and hole filling) lies on the scene surface: more than 10,000 points, median
    distance to the floor-and-boxes surface below 2 cm, fewer than 2 % farther than 10 cm."""
    from cumulus3d_amd import fusion, neighbors

    sc, src, ranges = synth
    be = pm.HipPatchMatch()
    h, w = sc.views[0].gray.shape
    nl = pm.num_levels(min(w, h), 3, 64)
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, ranges, pmp, fl, levels=[be.build_pyramid(v.gray, v.k, nl) for v in sc.views])
    views = list(range(len(sc.views)))
    with be.begin(inp) as s:
        outs, _, _, tm = s.densify(views, abi.Schedule(post=abi.PostParams.fast()))
    fp = fusion.FusionParams()
    count, q75 = neighbors.pair_stats(be, sc)
    overlap = [neighbors.select(count, q75, v, fp.consistency_num_images, 0.0) for v in views]
    keep = []
    arr = fusion.fusion_views(sc, [o.depth for o in outs], [o.normal for o in outs], [o.cost for o in outs], overlap, src, keep)
    cloud = fusion.fuse(be, arr, fp, int(sum((o.depth > 0).sum() for o in outs)) + 1)
    d = scenes.surface_distance(cloud.positions)
    n = len(d)
    print(f"{n} points, median surface distance {np.median(d):.4f} m, beyond 0.1 m {100.0 * (d > 0.1).mean():.3f} %, "
          f"post removed {tm.post_removed} filled {tm.post_filled}")
    assert n > 10_000
    assert np.median(d) < 0.02
    assert (d > 0.1).sum() * 50 < n
    be.close()
