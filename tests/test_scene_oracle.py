"""Pins of the scene-preparation oracle (oracle/scene_oracle.c: PairStats::new / select and
depth_ranges, crates/dense/src/neighbors.rs:47-189), independent of its restatement:

  * the reference's own tests: host_parts.rs:61-89 (selection orders with and without the angle
    bound, the range 7.5..12.5) and neighbors.rs:193-198 (in test_oracle_kats.py);
  * a brute-force recomputation in Python (math.acos, numpy) of counts, 75th-percentile angles,
    selections and ranges on seeded random tracks with repeated views and points behind cameras.

The host selection of the product library (c3d_select_views / c3d_select_diverse, no device) is
checked here too.
"""
import math

import numpy as np

import oracle
from cumulus3d_amd import neighbors, scenes


def bare_view(c):
    # a camera looking down from c (host_parts.rs:41-58)
    r = np.array([[1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, -1.0]])
    t = np.array([-(r[0, 0] * c[0]), -(r[1, 1] * c[1]), -(r[2, 2] * c[2])])
    return scenes.DenseView("", 4, 4, np.array([100.0, 100.0, 1.5, 1.5]), r, t, np.zeros((4, 4), np.uint8))


def reference_scene():
    views = [bare_view([0.0, 0.0, 10.0]), bare_view([3.0, 0.0, 10.0]), bare_view([0.05, 0.0, 10.0]), bare_view([-3.0, 0.0, 10.0])]
    pts = []
    for n, a, b in [(5, 0, 1), (8, 0, 2), (3, 0, 3), (1, 1, 3)]:
        pts += [scenes.ScenePoint(np.array([k * 0.1, 0.0, 0.0]), [a, b]) for k in range(n)]
    return scenes.DenseScene(views, pts, None, None)


def test_reference_neighbor_selection():
    sc = reference_scene()
    count, q75 = oracle.pair_stats(sc)
    a1 = math.radians(1.0)
    assert oracle.select(count, q75, 0, 10, a1) == [1, 3]
    assert oracle.select(count, q75, 0, 1, a1) == [1]
    assert oracle.select(count, q75, 0, 10, 0.0) == [2, 1, 3]
    assert oracle.select(count, q75, 1, 10, a1) == [0, 3]
    r = oracle.depth_ranges(sc)
    assert abs(r[0][0] - 7.5) < 1e-9 and abs(r[0][1] - 12.5) < 1e-9
    # the library's host selection on the same statistics
    assert neighbors.select(count, q75, 0, 10, a1) == [1, 3]
    assert neighbors.select(count, q75, 0, 1, a1) == [1]
    assert neighbors.select(count, q75, 0, 10, 0.0) == [2, 1, 3]
    assert neighbors.select(count, q75, 1, 10, a1) == [0, 3]
    # direction diversity off (decay 1) is the plain selection
    assert neighbors.select_diverse(sc, count, q75, 0, 10, a1, 8, 1.0) == [1, 3]


def random_tracks(nviews, npts, seed):
    rng = np.random.default_rng(seed)
    views = []
    for v in range(nviews):
        c = np.array([rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(8, 30)])
        views.append(bare_view(c))
    pts = []
    for _ in range(npts):
        x = np.array([rng.uniform(-25, 25), rng.uniform(-25, 25), rng.uniform(-3, 3)])
        if rng.random() < 0.03:
            x[2] = 40.0  # above every camera: behind all of them
        m = int(rng.integers(0, 7))
        tr = [int(q) for q in rng.integers(0, nviews, m)]
        if m >= 2 and rng.random() < 0.1:
            tr.append(tr[0])  # a view listed twice
        pts.append(scenes.ScenePoint(x, tr))
    return scenes.DenseScene(views, pts, None, None)


def brute(sc):
    n = len(sc.views)
    centers = [v.center() for v in sc.views]
    angs = {}
    depths = [[] for _ in range(n)]
    for p in sc.points:
        for i in range(len(p.views)):
            for j in range(i):
                a, b = p.views[i], p.views[j]
                if a == b:
                    continue
                u, w = centers[a] - p.xyz, centers[b] - p.xyz
                nu, nw = math.sqrt(u @ u), math.sqrt(w @ w)
                cc = min(1.0, max(-1.0, float(u @ w) / (nu * nw)))
                t = math.acos(cc)
                angs.setdefault((min(a, b), max(a, b)), []).append(float(np.float32(min(t, math.pi - t))))
        for v in p.views:
            d = float(sc.views[v].r[2] @ p.xyz + sc.views[v].t[2])
            if d > 0:
                depths[v].append(d)
    count = np.zeros((n, n), np.uint32)
    q75 = np.zeros((n, n))
    for (a, b), l in angs.items():
        l = sorted(l)
        pos = 0.75 * (len(l) - 1)
        lo = int(math.floor(pos))
        hi = min(lo + 1, len(l) - 1)
        q = l[lo] + (l[hi] - l[lo]) * (pos - lo)
        count[a, b] = count[b, a] = len(l)
        q75[a, b] = q75[b, a] = q
    ranges = []
    for l in depths:
        if not l:
            ranges.append(None)
            continue
        l = sorted(l)
        m = len(l)
        ranges.append((0.75 * l[min(int(math.floor(m * 0.01)), m - 1)], 1.25 * l[min(int(math.floor(m * 0.99)), m - 1)]))
    return count, q75, ranges


def test_pair_stats_and_ranges_match_brute_force():
    for (nv, npts, seed) in [(6, 400, 1), (13, 2000, 2), (3, 5, 3)]:
        sc = random_tracks(nv, npts, seed)
        count, q75 = oracle.pair_stats(sc)
        bc, bq, br = brute(sc)
        assert np.array_equal(count, bc)
        assert np.allclose(q75, bq, rtol=1e-12, atol=0)
        assert oracle.depth_ranges(sc) == br
        a = math.radians(5.0)
        for r in range(nv):
            for mx in (1, 3, 50):
                want = sorted([o for o in range(nv) if o != r and bc[r, o] > 0 and bq[r, o] >= a], key=lambda o: (-int(bc[r, o]), o))
                assert oracle.select(count, q75, r, mx, a) == want[:mx]
                assert neighbors.select(count, q75, r, mx, a) == want[:mx]


def brute_diverse(sc, count, q75, r, mx, amin, bins, decay):
    n = len(sc.views)
    cands = sorted([o for o in range(n) if o != r and count[r, o] > 0 and not (amin > 0 and q75[r, o] < amin)],
                   key=lambda o: (-int(count[r, o]), o))
    vr = sc.views[r]
    info = []
    for o in cands:
        b = vr.r @ (sc.views[o].center() - vr.center())
        lat = math.sqrt(b[0] ** 2 + b[1] ** 2)
        if lat < 0.3 * abs(b[2]):
            bin_ = bins
        else:
            bin_ = min(int((math.atan2(b[1], b[0]) + math.pi) / (2 * math.pi) * bins), bins - 1)
        info.append((o, float(count[r, o]), bin_))
    used = [0] * (bins + 1)
    taken = [False] * len(info)
    out = []
    while len(out) < mx:
        best = None
        for i, (o, s, bn) in enumerate(info):
            if taken[i]:
                continue
            v = s * decay ** used[bn]
            if best is None or v > best[0]:
                best = (v, i)
        if best is None:
            break
        taken[best[1]] = True
        used[info[best[1]][2]] += 1
        out.append(info[best[1]][0])
    return out


def test_select_diverse_matches_brute_force():
    sc = random_tracks(14, 3000, 5)
    count, q75 = oracle.pair_stats(sc)
    for r in range(14):
        for (bins, decay, mx) in [(8, 0.5, 6), (4, 0.25, 10), (1, 0.5, 3), (8, 1.0, 5)]:
            want = brute_diverse(sc, count, q75, r, mx, math.radians(2.0), bins, decay)
            assert neighbors.select_diverse(sc, count, q75, r, mx, math.radians(2.0), bins, decay) == want
