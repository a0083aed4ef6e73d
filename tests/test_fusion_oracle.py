"""Pins of the agreement-fusion oracle (oracle/fusion_oracle.c, fuse_consistency,
crates/dense/src/fusion.rs:202-490) and of the cloud comparison, CPU only:

* exact depth maps of the plane z = 0 seen by a formation of cameras fuse into points on the plane
  (|z| < 1e-5 m) with normals (0, 0, +-1) and keys in the reference's order (views, rows,
  columns); asking for more agreeing views than there are overlap views gives no point, and
  without mark_used there are more points;
* a view whose depth is 5 % off agrees with nobody at a 0.5 % tolerance;
* the cloud comparison of a cloud and its copy shifted by 1 cm is 1 cm with full completeness at
  5 cm, and 10 cm away it is empty.
"""
import numpy as np

from cumulus3d_amd import fusion, scenes


def _plane_scene(w=64, h=48, stations=4):
    cams = scenes.formation_rig(stations, 3, spacing=4.0)
    views, depth, normal = [], [], []
    for i, (c, target) in enumerate(cams):
        r, t = scenes.look_at(c, target)
        k = np.array([60.0, 60.0, (w - 1) / 2.0, (h - 1) / 2.0])
        ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
        rays = np.stack([(xs - k[2]) / k[0], (ys - k[3]) / k[1], np.ones_like(xs)], -1)  # camera
        dw = rays @ r  # world directions (r^T ray)
        s = -c[2] / dw[..., 2]  # ray parameter to z = 0 (camera depth equals s since ray z = 1)
        d = np.where(s > 0, s, 0.0).astype(np.float32)
        n_cam = np.broadcast_to(r @ np.array([0.0, 0.0, 1.0]), (h, w, 3)).astype(np.float32)
        gray = (np.arange(w * h) % 251).astype(np.uint8).reshape(h, w)
        views.append(scenes.DenseView(f"p{i}", w, h, k, r, t, gray))
        depth.append(d)
        normal.append(np.ascontiguousarray(n_cam))
    sc = scenes.DenseScene(views, [], None, None)
    return sc, depth, normal


def test_plane_fuses_onto_the_plane():
    import oracle

    sc, depth, normal = _plane_scene()
    n = len(sc.views)
    src = scenes.nearest_sources(sc, [4] * n)
    keep = []
    views = fusion.fusion_views(sc, depth, normal, [np.full_like(d, 0.1) for d in depth], src, src, keep)
    cloud = oracle.fuse_consistency(views, fusion.FusionParams(min_consistent_views=2), 10 ** 6)
    assert len(cloud.positions) > 1000
    assert np.abs(cloud.positions[:, 2]).max() < 1e-5
    assert np.allclose(np.abs(cloud.normals[:, 2]), 1.0, atol=1e-6)
    assert np.all(np.diff(cloud.keys.astype(np.int64)) > 0)
    # more agreeing views than overlap views -> nothing; without mark_used every agreeing pixel of a
    # later view is also a reference pixel -> more points
    none = oracle.fuse_consistency(views, fusion.FusionParams(min_consistent_views=5), 10 ** 6)
    assert len(none.positions) == 0
    more = oracle.fuse_consistency(views, fusion.FusionParams(min_consistent_views=2, mark_used=False), 10 ** 6)
    assert len(more.positions) > len(cloud.positions)
    # a reference pixel whose depth is 5 % off finds no agreement at 0.5 %
    bad = [d * 1.05 if i == 0 else d for i, d in enumerate(depth)]
    views2 = fusion.fusion_views(sc, bad, normal, None, src, src, keep)
    c2 = oracle.fuse_consistency(views2, fusion.FusionParams(min_consistent_views=1), 10 ** 6)
    assert not np.any((c2.keys >> np.uint64(32)) == 0)


def test_cloud_compare_closed_form():
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 10, (5000, 3)).astype(np.float32)
    a = fusion.Cloud(pts, pts, np.zeros((5000, 3), np.uint8), np.arange(5000, dtype=np.uint64))
    b = fusion.Cloud(pts + np.float32([0.01, 0, 0]), pts, a.colors, a.keys)
    med, comp = fusion.compare(a, b, 0.05)
    assert abs(med - 0.01) < 1e-4 and comp == 1.0
    far = fusion.Cloud(pts + np.float32([0, 0, 0.1]) + 100, pts, a.colors, a.keys)
    med, comp = fusion.compare(a, far, 0.05)
    assert comp == 0.0
