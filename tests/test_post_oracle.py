"""Pins of the post-processing oracle (oracle/post_oracle.c: remove_speckles and fill_holes,
crates/dense/src/postproc.rs:77-204), independent of its restatement:

  * the reference's own test (postproc.rs:208-231): a 2x2 speckle erased (4 pixels), a 3x3 hole
    and the erased speckle filled (13 pixels), the hole at depth 10;
  * components against scipy's connected components of the join graph built with numpy, sizes and
    medians against numpy (so a wrong neighbour, join rule, median index or area bound fails);
  * fills of a tilted plane against the plane's closed-form depth, the gray-edge stop, the raw-cost
    bound and the least-gray-difference choice.
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import oracle
from cumulus3d_amd import abi


def test_reference_speckle_and_fill():
    w, h = 40, 30
    k = [50.0, 50.0, 19.5, 14.5]
    d = np.full((h, w), 10.0, np.float32)
    n = np.zeros((h, w, 3), np.float32)
    n[..., 2] = -1.0
    c = np.full((h, w), 0.5, np.float32)
    for (x, y) in [(5, 5), (6, 5), (5, 6), (6, 6)]:
        d[y, x] = 5.0
    d[20:23, 20:23] = 0.0
    p = abi.PostParams(speckle_min_area_frac=10.0 / (w * h), fill_holes=True, fill_max_frac=0.1)
    assert oracle.remove_speckles(d, n, c, p) == 4
    assert d[5, 5] == 0.0
    gray = np.full((h, w), 100, np.uint8)
    raw = np.full((h, w), 0.3, np.float32)
    assert oracle.fill_holes(d, n, c, raw, gray, k, p) == 13
    assert abs(d[21, 21] - 10.0) < 1e-4


def speckle_map(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    d = (20.0 + 0.02 * xx + 0.01 * yy).astype(np.float32)
    d *= (1.0 + 0.003 * rng.standard_normal((h, w))).astype(np.float32)
    # blobs at other depths, invalid pixels, isolated spikes
    for _ in range(max(1, h * w // 300)):
        y0, x0 = rng.integers(0, h), rng.integers(0, w)
        r = rng.integers(1, 5)
        d[max(0, y0 - r):y0 + r, max(0, x0 - r):x0 + r] = rng.uniform(5.0, 60.0)
    d[rng.random((h, w)) < 0.08] = 0.0
    d[rng.random((h, w)) < 0.02] = -1.0
    spikes = rng.random((h, w)) < 0.02
    d[spikes] = rng.uniform(1.0, 100.0, spikes.sum())
    nrm = rng.standard_normal((h, w, 3)).astype(np.float32)
    cost = rng.uniform(0.0, 1.0, (h, w)).astype(np.float32)
    cost[rng.random((h, w)) < 0.3] = np.float32(0.1)  # ties in the medians
    return d.astype(np.float32), nrm, cost


def expected_speckles(d, cost, p):
    """Independent recomputation: components of the join graph (scipy), erase rule (numpy)."""
    h, w = d.shape
    min_area = int(np.ceil(h * w * p.speckle_min_area_frac))
    valid = ~(d <= 0)
    idx = np.arange(h * w).reshape(h, w)
    rows, cols = [], []
    rel = np.float32(p.speckle_rel_depth)
    for a, b in (((slice(None), slice(0, -1)), (slice(None), slice(1, None))), ((slice(0, -1), slice(None)), (slice(1, None), slice(None)))):
        da, db = d[a], d[b]
        join = valid[a] & valid[b] & (np.abs(db - da) <= rel * np.minimum(da, db))
        rows.append(idx[a][join])
        cols.append(idx[b][join])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    g = sp.coo_matrix((np.ones(len(r)), (r, c)), shape=(h * w, h * w))
    _, lab = connected_components(g, directed=False)
    lab = lab.reshape(h, w)
    erase = np.zeros((h, w), bool)
    if min_area <= 1:
        return erase
    for L in np.unique(lab[valid]):
        m = (lab == L) & valid
        if m.sum() < min_area:
            cs = np.sort(cost[m])
            if cs[len(cs) // 2] > np.float32(p.speckle_keep_cost):
                erase |= m
    return erase


def test_speckles_match_independent_components():
    for (h, w, seed, frac) in [(30, 41, 1, 2e-2), (64, 64, 2, 5e-3), (17, 90, 3, 1e-2), (1, 200, 4, 3e-2), (50, 1, 5, 1e-1)]:
        d, nrm, cost = speckle_map(h, w, seed)
        p = abi.PostParams(speckle_min_area_frac=frac)
        want = expected_speckles(d, cost, p)
        d1, n1 = d.copy(), nrm.copy()
        removed = oracle.remove_speckles(d1, n1, cost, p)
        assert removed == want.sum()
        assert np.array_equal(d1 == 0, (d == 0) | want)
        assert np.array_equal(d1[~want], d[~want]) and np.all(n1[want] == 0)
        assert want.sum() > 0


def test_speckles_nothing_below_area_two():
    d, nrm, cost = speckle_map(20, 20, 7)
    p = abi.PostParams(speckle_min_area_frac=1.0 / 400)  # ceil(400 / 400) = 1: no removal
    d1 = d.copy()
    assert oracle.remove_speckles(d1, nrm.copy(), cost, p) == 0
    assert np.array_equal(d1, d)


def plane_depth(k, n, c0, h, w):
    """Depth of the plane n . X = c0 (camera coordinates) along each pixel's ray."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ray = np.stack([(xx - k[2]) / k[0], (yy - k[3]) / k[1], np.ones_like(xx)], -1)
    return c0 / (ray @ n)


def test_fill_reproduces_a_tilted_plane():
    h, w = 48, 64
    k = [60.0, 62.0, 31.5, 23.5]
    n = np.array([0.2, -0.3, -1.0])
    n /= np.linalg.norm(n)
    dd = plane_depth(k, n, -25.0, h, w)
    d = dd.astype(np.float32)
    nrm = np.broadcast_to(n.astype(np.float32), (h, w, 3)).copy()
    rng = np.random.default_rng(3)
    holes = rng.random((h, w)) < 0.2
    d[holes] = 0.0
    cost = rng.uniform(0.0, 0.5, (h, w)).astype(np.float32)
    raw = cost.copy()
    raw[5, :] = 2.0  # above fill_max_cost: never filled
    gray = np.full((h, w), 80, np.uint8)
    p = abi.PostParams(fill_holes=True, fill_max_frac=0.1)
    d1, n1, c1 = d.copy(), nrm.copy(), cost.copy()
    filled = oracle.fill_holes(d1, n1, c1, raw, gray, k, p)
    blocked = holes & (raw > 1.0)
    got = holes & ~blocked & (d1 > 0)
    assert filled == got.sum() and got.sum() > 0.9 * (holes & ~blocked).sum()
    assert np.all(d1[blocked] == 0)
    assert np.allclose(d1[got], dd[got], rtol=2e-6)
    assert np.array_equal(n1[got], nrm[got])
    # the copied cost is one of the valid neighbours'
    assert np.all(np.isin(c1[got], cost[~holes]))


def test_fill_stops_at_gray_edges_and_prefers_similar_gray():
    h, w = 9, 21
    k = [30.0, 30.0, 10.0, 4.0]
    d = np.zeros((h, w), np.float32)
    d[:, :8] = 10.0
    d[:, 13:] = 20.0
    nrm = np.zeros((h, w, 3), np.float32)
    nrm[..., 2] = -1.0
    cost = np.full((h, w), 0.2, np.float32)
    gray = np.zeros((h, w), np.uint8)
    gray[:, :10] = 50
    gray[:, 10:] = 200  # an edge between columns 9 and 10 (a step of 150 > 15)
    p = abi.PostParams(fill_holes=True, fill_max_frac=0.5)
    d1 = d.copy()
    oracle.fill_holes(d1, nrm.copy(), cost.copy(), cost.copy(), gray, k, p)
    # fronto-parallel planes: depths are copied; each side fills from its own gray region
    assert np.all(d1[:, 8:10] == 10.0) and np.all(d1[:, 10:13] == 20.0)
    # with the edge removed, a hole pixel takes the nearer valid pixel; at equal distance (column
    # 10: 3 to both sides) the first direction of the reference's list, +x
    gray[:] = 50
    d2 = d.copy()
    oracle.fill_holes(d2, nrm.copy(), cost.copy(), cost.copy(), gray, k, p)
    assert np.all(d2[:, 8:10] == 10.0) and np.all(d2[:, 10:13] == 20.0)
    # a smaller gray difference beats a shorter walk
    gray[:, 11:] = 60
    d3 = d.copy()
    oracle.fill_holes(d3, nrm.copy(), cost.copy(), cost.copy(), gray, k, p)
    assert np.all(d3[:, 10] == 10.0)


def test_fill_tie_takes_earlier_direction():
    # a hole pixel with valid pixels at the same distance and gray difference left and right: the
    # first direction of the reference's list (+x) wins
    h, w = 1, 7
    k = [10.0, 10.0, 3.0, 0.0]
    d = np.array([[0, 0, 5.0, 0, 7.0, 0, 0]], np.float32)
    nrm = np.zeros((h, w, 3), np.float32)
    nrm[..., 2] = -1.0
    cost = np.zeros((h, w), np.float32)
    gray = np.full((h, w), 9, np.uint8)
    p = abi.PostParams(fill_holes=True, fill_max_frac=1.0)
    d1 = d.copy()
    oracle.fill_holes(d1, nrm.copy(), cost.copy(), cost.copy(), gray, k, p)
    assert d1[0, 3] == 7.0  # +x at distance 1 before -x at distance 1
    assert d1[0, 1] == 5.0 and d1[0, 0] == 5.0 and d1[0, 5] == 7.0 and d1[0, 6] == 7.0
