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


def _two_down_views(hA, hB, w=64, h=48):
    """Two cameras looking straight down at the plane z = 0 from heights hA, hB: every depth map is
    the constant camera height, every normal (0, 0, -1) in camera coordinates."""
    k = np.array([60.0, 60.0, 31.5, 23.5])
    r = np.diag([1.0, -1.0, -1.0])
    views = []
    for c, g in (((0.0, 0.0, hA), 100), ((0.7, 0.3, hB), 200)):
        c = np.array(c)
        views.append(scenes.DenseView(f"v{g}", w, h, k, r, -r @ c, np.full((h, w), g, np.uint8)))
    return scenes.DenseScene(views, [], None, None), k, r


def test_inverse_variance_weights_follow_the_fourth_power_law():
    """Fused point of a pixel of view 0 and its one agreeing pixel of view 1 against the closed
    form of inverse-variance weighting: sigma_d = d^2 sigma_px / (f b) gives w ~ 1 / (d^4 sigma_px^2)
    (baselines and focal lengths equal here), sigma_px = sigma0 + slope * clamp(cost, 0, 2); the
    fused point and colour divide the two observations in the ratio w1 : w0."""
    import oracle

    w, h = 64, 48
    for (hA, hB, cB, inv) in [(10.0, 10.0, 0.0, True), (10.0, 12.0, 0.0, True), (10.0, 10.0, 0.25, True), (10.0, 10.0, 5.0, True),
                              (10.0, 12.0, 0.25, False)]:
        sc, k, r = _two_down_views(hA, hB, w, h)
        depth = [np.full((h, w), hA, np.float32), np.full((h, w), hB, np.float32)]
        normal = [np.broadcast_to(np.array([0.0, 0.0, -1.0], np.float32), (h, w, 3)).copy() for _ in range(2)]
        cost = [np.zeros((h, w), np.float32), np.full((h, w), cB, np.float32)]
        keep = []
        arr = fusion.fusion_views(sc, depth, normal, cost, [[1], [0]], [[1], [0]], keep)
        p = fusion.FusionParams(min_consistent_views=1, consistency_num_images=1, mark_used=False, inverse_variance=inv)
        cl = oracle.fuse_consistency(arr, p, 2 * w * h + 1)
        sA, sB = 0.25, 0.25 + 1.0 * min(max(cB, 0.0), 2.0)
        wA, wB = (1.0, 1.0) if not inv else (1.0 / (hA ** 4 * sA ** 2), 1.0 / (hB ** 4 * sB ** 2))
        t = wB / (wA + wB)
        cA, cBc = sc.views[0].center(), sc.views[1].center()
        checked = 0
        for (row, col) in [(5, 7), (23, 31), (40, 60)]:
            ray = np.array([(col - k[2]) / k[0], (row - k[3]) / k[1], 1.0])
            xA = cA + (r.T @ ray) * hA                        # the pixel's point on the plane
            cam = r @ (xA - cBc)                              # in view 1
            u, v = k[0] * cam[0] / cam[2] + k[2], k[1] * cam[1] / cam[2] + k[3]
            qc, qr = np.floor(u + 0.5), np.floor(v + 0.5)
            if not (0 <= qc < w and 0 <= qr < h):
                continue
            rq = np.array([(qc - k[2]) / k[0], (qr - k[3]) / k[1], 1.0])
            xB = cBc + (r.T @ rq) * hB
            want = xA + t * (xB - xA)
            key = (row << 16) | col
            i = np.nonzero(cl.keys == key)[0]
            assert len(i) == 1
            assert np.allclose(cl.positions[i[0]], want, atol=2e-5), (hA, hB, cB, inv)
            assert abs(int(cl.colors[i[0], 0]) - (100 + 100 * t)) <= 0.5 + 1e-9
            checked += 1
        assert checked >= 2
