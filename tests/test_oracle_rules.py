"""Known-answer tests of the oracle's decision rules: propagation region pick and plane transfer,
view selection and aggregation, refinement acceptance, the normal perturbation's retries, the
top-K mean and the read-out tests. Every case is derived by hand from the rule text of
crates/dense/src/kernel.rs:50-66 (and, for the offset order, the top-K and the read-out, from
crates/cuda/src/kernels/patchmatch.cu:327-345,425-436,714-733 and densify.rs:446-454), with its
arithmetic done independently in Python; the oracle's half step, run start and read-out call these
functions with the costs they compute, so a slip in any of them fails a case here.

Constants of the cases: tau(t) = 0.75 and tau1 = 1.25 (both exact in float32, so costs equal to
them test the strict comparisons), beta = 0.3, prev_view_weight = 0.2, lambda = 0.2, n1 = 2,
n2 = 3 (the reference's PmParams defaults, params.rs:128-154, apart from tau0 and tau1).
"""
import math

import numpy as np
import pytest

import oracle

TAU, TAU1, BETA, PREV_W, LAM, N1, N2 = 0.75, 1.25, 0.3, 0.2, 0.2, 2, 3


def _params(o):
    return o.SelectParams(TAU, TAU1, 1.0 / (2.0 * BETA * BETA), PREV_W, LAM, N1, N2)


def _conf(costs):
    """Weight of an adopted view: the mean of exp(-m^2 / (2 beta^2)) over its good costs (kernel.rs:55)."""
    return sum(math.exp(-c * c / (2 * BETA * BETA)) for c in costs) / len(costs)


def _table(current, propagated):
    """One source's row: hypothesis 0 is the current plane, 1-8 the propagated ones."""
    return np.array([current] + list(propagated), np.float32)


# ---------------------------------------------------------------- propagation (kernel.rs:50-52)

def test_region_pick_offsets_order_and_ties(o=oracle):
    """Up V: offsets (0,-1),(-1,-2),(1,-2),(-2,-3),(2,-3),(-3,-4),(3,-4) in that order
    (patchmatch.cu:425); the other directions are 90-degree rotations up -> right -> down -> left
    ((dx, dy) -> (-dy, dx) with y pointing down); bands (0,-3),(0,-5),...,(0,-23). The smallest stored
    cost wins, the first offset in order on ties; offsets outside the image are skipped."""
    w = h = 61
    x = y = 30
    base = np.ones((h, w), np.float32)

    c = base.copy()
    c[y - 3, x - 2] = 0.25  # (-2,-3), 4th V offset
    c[y - 3, x + 2] = 0.25  # (2,-3), 5th: same cost, later
    assert o.region_pick(c, x, y, 0) == (x - 2, y - 3)
    c[y - 4, x + 3] = 0.125  # strictly smaller, last offset
    assert o.region_pick(c, x, y, 0) == (x + 3, y - 4)

    # right V: (0,-1) -> (1,0), (-1,-2) -> (2,-1), (1,-2) -> (2,1)
    c = base.copy()
    c[y + 1, x + 2] = 0.5
    c[y - 1, x + 2] = 0.5
    assert o.region_pick(c, x, y, 1) == (x + 2, y - 1)
    # down V: (0,-1) -> (0,1) after two rotations; left V: (0,-1) -> (-1,0)
    c = base.copy()
    c[y + 1, x] = 0.5
    c[y, x - 1] = 0.5
    assert o.region_pick(c, x, y, 2) == (x, y + 1)
    assert o.region_pick(c, x, y, 3) == (x - 1, y)
    # bands: up (0,-3)...(0,-23); the nearer of two equal costs wins; right band (3,0)...(23,0)
    c = base.copy()
    c[y - 23, x] = 0.5
    c[y - 7, x] = 0.5
    assert o.region_pick(c, x, y, 4) == (x, y - 7)
    c[y, x + 23] = 0.0
    assert o.region_pick(c, x, y, 5) == (x + 23, y)
    assert o.region_pick(c, x, y, 6) == (x, y + 3)  # all equal: the first offset
    assert o.region_pick(c, x, y, 7) == (x - 3, y)
    # the V shape does not reach (0,-3): a cheaper band pixel is not picked by region 0
    c = base.copy()
    c[y - 3, x] = 0.0
    assert o.region_pick(c, x, y, 0) == (x, y - 1)
    # at the image border: offsets outside are skipped, none inside -> nothing
    c = base.copy()
    assert o.region_pick(c, 0, 0, 0) is None
    assert o.region_pick(c, 0, 0, 3) is None
    assert o.region_pick(c, 0, 0, 4) is None
    assert o.region_pick(c, 1, 2, 0) == (1, 1)
    c[0, 2] = 0.5  # (1,-2) from (1, 2)
    assert o.region_pick(c, 1, 2, 0) == (2, 0)


def test_transfer_plane_onto_own_ray(o=oracle):
    """The neighbour's plane moved onto the pixel's ray is the same 3-D plane: the point d * ray lies
    on the plane through d_q * ray_q with normal n (kernel.rs:51). Refused when the plane does not
    face the camera along the ray (n . ray >= -1e-6), the neighbour has no depth, or the
    intersection is behind the camera."""
    # fronto-parallel plane z = 10
    out = o.transfer_plane([0, 0, -1, 10.0], [0.1, 0.0, 1.0], [0.2, 0.1, 1.0])
    assert out is not None and abs(out[3] - 10.0) < 1e-5 and list(out[:3]) == [0, 0, -1]
    # tilted plane through P = d_q ray_q: depth = (n . P) / (n . ray)
    n = np.array([0.3, 0.1, -1.0])
    n /= np.linalg.norm(n)
    rq, r = np.array([0.1, -0.05, 1.0]), np.array([-0.2, 0.15, 1.0])
    P = 12.0 * rq
    want = float(n @ P / (n @ r))
    out = o.transfer_plane([*n, 12.0], rq, r)
    assert out is not None and abs(out[3] - want) < 1e-5 * want
    assert np.allclose(out[:3], n.astype(np.float32))
    # facing away, edge-on, no depth, behind the camera
    assert o.transfer_plane([0, 0, 1, 10.0], [0, 0, 1], [0, 0, 1]) is None
    assert o.transfer_plane([1, 0, 0, 10.0], [0, 0, 1], [0, 0, 1]) is None
    assert o.transfer_plane([0, 0, -1, 0.0], [0, 0, 1], [0, 0, 1]) is None
    m = np.array([1.0, 0.0, -0.1])
    m /= np.linalg.norm(m)  # n . ray_q > 0 at ray_q = (1, 0, 1)/.. while n . ray < 0 at ray = (-1, 0, 1)
    assert o.transfer_plane([*m, 5.0], [1.0, 0.0, 1.0], [-1.0, 0.0, 1.0]) is None


# ---------------------------------------------------------------- view selection (kernel.rs:53-58)

def test_vote_edges(o=oracle):
    """A view is adopted when its good count (m < tau) exceeds n1 and its bad count (m > tau1) is
    below n2, counted over the 8 propagated hypotheses only. Costs equal to tau are not good and
    costs equal to tau1 are not bad."""
    p = _params(o)
    neutral = 1.0  # neither good nor bad

    def adopted(row, vprev=255):
        _, _, _, _, view = o.select_views(p, [1.0], [row], vprev=vprev)
        return view == 0

    assert adopted(_table(neutral, [0.0] * 3 + [neutral] * 5))          # 3 good > 2
    assert not adopted(_table(neutral, [0.0] * 2 + [neutral] * 6))      # 2 good = n1
    assert not adopted(_table(0.0, [0.0] * 2 + [neutral] * 6))          # the current plane does not vote
    assert not adopted(_table(neutral, [0.0] * 2 + [TAU] + [neutral] * 5))  # m = tau is not good
    assert adopted(_table(neutral, [0.0] * 3 + [2.0] * 2 + [neutral] * 3))  # 2 bad < 3
    assert not adopted(_table(neutral, [0.0] * 3 + [2.0] * 3 + [neutral] * 2))  # 3 bad = n2
    assert adopted(_table(neutral, [0.0] * 3 + [2.0] * 2 + [TAU1] * 3))  # m = tau1 is not bad


def test_view_weight_previous_view_and_prior(o=oracle):
    """Adopted: w = mean exp(-m^2/(2 beta^2)) over the good costs, doubled for the previous most
    important view; not adopted: prev_view_weight for the previous view, else 0; w'' = w x prior."""
    p = _params(o)
    good = [0.1, 0.2, 0.3]
    row = _table(1.0, good + [1.0] * 5)
    w = _conf(good)
    wv, den, _, _, view = o.select_views(p, [1.0], [row])
    assert view == 0 and wv[0] == pytest.approx(w, rel=1e-6) and den == pytest.approx(w, rel=1e-6)
    wv, _, _, _, _ = o.select_views(p, [1.0], [row], vprev=0)
    assert wv[0] == pytest.approx(2 * w, rel=1e-6)
    wv, _, _, _, _ = o.select_views(p, [0.5], [row], vprev=0)
    assert wv[0] == pytest.approx(2 * w * 0.5, rel=1e-6)
    # not adopted but the previous view: prev_view_weight x prior, and it takes part
    lost = _table(1.0, [0.1] * 2 + [1.0] * 6)
    wv, den, agg, _, view = o.select_views(p, [0.5, 1.0], [lost, lost], vprev=0)
    assert wv[0] == pytest.approx(PREV_W * 0.5, rel=1e-6) and den == pytest.approx(PREV_W * 0.5, rel=1e-6)
    assert view == 0  # the most important view is the one with the largest positive w''
    assert wv[1] <= 0  # the other view takes no part
    # previous view index past the sources: nobody is the previous view
    wv, den, _, _, view = o.select_views(p, [1.0], [row], vprev=255)
    assert wv[0] == pytest.approx(w, rel=1e-6)


def test_weighted_aggregation_and_hypothesis_choice(o=oracle):
    """The adopted views' weighted mean of m (+ lambda g in a geometric run) per hypothesis; the
    minimum is adopted."""
    p = _params(o)
    g1 = [0.0, 0.1, 0.2]
    g2 = [0.3, 0.3, 0.4, 0.2]
    rows = [_table(0.9, [0.5, 0.1, 0.2, 0.0, 1.0, 1.0, 0.4, 0.6]),
            _table(0.3, [0.6, 0.3, 0.3, 0.4, 0.2, 1.0, 1.0, 1.0])]
    # view 0 good: 0.5, 0.1, 0.2, 0.0, 0.4, 0.6 (6 good, 0 bad); view 1 good: 0.6, 0.3, 0.3, 0.4, 0.2
    w = [_conf([0.5, 0.1, 0.2, 0.0, 0.4, 0.6]) * 0.7, _conf([0.6, 0.3, 0.3, 0.4, 0.2]) * 0.9]
    wv, den, agg, best, view = o.select_views(p, [0.7, 0.9], rows)
    assert np.allclose(wv, w, rtol=1e-6) and den == pytest.approx(sum(w), rel=1e-6)
    want = (w[0] * rows[0].astype(float) + w[1] * rows[1].astype(float)) / sum(w)
    assert np.allclose(agg, want, rtol=1e-5)
    assert best == int(np.argmin(want)) and view == int(np.argmax(w))
    assert g1 and g2  # the geometric case below
    geo = [np.array([0.0, 0.5, 3.0, 0.0, 1.0, 0.0, 0.0, 2.0, 0.0], np.float32),
           np.array([0.0, 0.0, 0.0, 0.0, 0.0, 3.0, 0.0, 0.0, 1.0], np.float32)]
    _, _, agg_g, best_g, _ = o.select_views(p, [0.7, 0.9], rows, geo)
    want_g = (w[0] * (rows[0] + LAM * geo[0]).astype(float) + w[1] * (rows[1] + LAM * geo[1]).astype(float)) / sum(w)
    assert np.allclose(agg_g, want_g, rtol=1e-5) and best_g == int(np.argmin(want_g))


def test_hypothesis_and_view_ties_go_to_the_earlier_index(o=oracle):
    p = _params(o)
    row = _table(0.5, [0.5, 0.1, 0.5, 0.5, 0.1, 0.5, 0.5, 0.5])  # hypotheses 2 and 5 tie at the minimum
    _, _, _, best, _ = o.select_views(p, [1.0], [row])
    assert best == 2
    row = _table(0.1, [0.5, 0.1, 0.1, 0.5, 0.1, 0.5, 0.5, 0.5])  # the current plane ties with 2, 3, 5
    _, _, _, best, _ = o.select_views(p, [1.0], [row])
    assert best == 0
    # two views with the same costs and priors: the same w''; the earlier view is the most important
    _, _, _, _, view = o.select_views(p, [1.0, 1.0, 1.0], [row, row, row])
    assert view == 0
    heavier = _table(0.1, [0.0, 0.0, 0.0, 0.5, 0.1, 0.5, 0.5, 0.5])
    _, _, _, _, view = o.select_views(p, [1.0, 1.0, 1.0], [row, heavier, heavier])
    assert view == 1


def test_prior_weighted_fallback(o=oracle):
    """No view adopted (and no previous view): the prior-weighted mean of m + lambda g over every
    view with a positive prior; refinement then uses the priors as weights; no most important view.
    No view with a positive prior: every aggregated cost is 2."""
    p = _params(o)
    rows = [_table(1.0, [0.5, 1.0, 0.7, 1.0, 1.0, 1.0, 1.0, 0.9]),   # 2 good only
            _table(0.6, [1.0, 0.5, 1.0, 1.0, 2.0, 2.0, 2.0, 1.0])]   # 3 bad
    pri = [0.25, 0.75]
    wv, den, agg, best, view = o.select_views(p, pri, rows)
    want = (pri[0] * rows[0].astype(float) + pri[1] * rows[1].astype(float)) / sum(pri)
    assert view == 255 and np.allclose(agg, want, rtol=1e-6) and best == int(np.argmin(want))
    assert np.allclose(wv, pri) and den == pytest.approx(sum(pri))
    wv, den, agg, best, view = o.select_views(p, [0.0, -1.0], rows)
    assert den == 0.0 and np.all(agg == 2.0) and best == 0 and view == 255 and np.all(wv == 0)


def test_views_with_zero_prior_are_not_evaluated(o=oracle):
    """A view whose prior is 0 takes no part: its costs are not read (NaN there changes nothing)."""
    p = _params(o)
    row = _table(0.4, [0.1, 0.2, 0.3, 1.0, 1.0, 1.0, 1.0, 1.0])
    junk = np.full(9, np.nan, np.float32)
    a = o.select_views(p, [1.0], [row])
    b = o.select_views(p, [1.0, 0.0], [row, junk])
    assert a[1] == b[1] and np.array_equal(a[2], b[2]) and a[3:] == b[3:] and b[0][1] == 0.0


# ---------------------------------------------------------------- refinement (kernel.rs:59-62)

def test_refinement_acceptance(o=oracle):
    """A candidate's cost is the weighted mean over the views taking part (wv > 0); it replaces the
    current best only when strictly smaller; candidates are tried in order."""
    p = _params(o)
    wv = [0.5, -0.25, 1.5]  # the middle view takes no part
    cand = np.array([[0.6, 100.0, 0.6],    # 0.6
                     [0.2, 100.0, 0.4],    # (0.1 + 0.6) / 2 = 0.35
                     [0.5, 0.0, 0.3],      # (0.25 + 0.45) / 2 = 0.35: ties with 1, not taken
                     [0.1, 0.0, 0.9]], np.float32)  # (0.05 + 1.35) / 2 = 0.7
    k, c = o.refine_accept(p, wv, 2.0, cand, bcost=0.5)
    assert k == 1 and c == pytest.approx(0.35, rel=1e-6)
    k, c = o.refine_accept(p, wv, 2.0, cand, bcost=0.35)  # equal to the best: kept
    assert k == -1 and c == pytest.approx(0.35, rel=1e-6)
    geo = np.zeros_like(cand)
    geo[1] = [1.0, 0.0, 1.0]  # candidate 1 becomes (0.5 (0.2 + 0.2) + 1.5 (0.4 + 0.2)) / 2 = 0.55
    k, c = o.refine_accept(p, wv, 2.0, cand, geo, bcost=0.5)
    assert k == 2 and c == pytest.approx(0.35, rel=1e-6)  # candidate 2 is then the first below 0.5
    # no weight at all: every candidate costs 2
    k, c = o.refine_accept(p, [-1.0, 0.0, -2.0], 0.0, cand, bcost=2.5)
    assert k == 0 and c == 2.0


def test_normal_perturbation_halves_up_to_four_tries(o=oracle):
    """n rotated by u phi about the axis a x n toward (a x n) x n; when the result does not face the
    camera the angle is halved, 4 tries in all, then n is kept (kernel.rs:60-61). With n = (0,0,-1)
    and a = (1,0,0) the rotation by theta gives (-sin theta, 0, -cos theta)."""
    n, a = [0.0, 0.0, -1.0], [1.0, 0.0, 0.0]
    rot = lambda t: np.array([-math.sin(t), 0.0, -math.cos(t)])  # noqa: E731
    out = o.perturb_normal_det(a, 1.0, n, [0, 0, 1.0], math.radians(60))
    assert np.allclose(out, rot(math.radians(60)), atol=1e-6)
    out = o.perturb_normal_det(a, 0.5, n, [0, 0, 1.0], math.radians(60))  # u scales the angle
    assert np.allclose(out, rot(math.radians(30)), atol=1e-6)
    # ray (-1, 0, 1): faces the camera while sin theta < cos theta (theta < 45 deg): 160 and 80 fail,
    # 40 is taken on the third try
    out = o.perturb_normal_det(a, 1.0, n, [-1.0, 0, 1.0], math.radians(160))
    assert np.allclose(out, rot(math.radians(40)), atol=1e-6)
    # an angle whose four halvings all fail keeps n
    faces = lambda t: -math.sin(t) * -1.0 + -math.cos(t) < 0.0  # noqa: E731
    t0 = next(t for t in np.arange(0.5, 40.0, 0.001) if not any(faces(t / 2 ** k) for k in range(4)) and faces(t / 16))
    assert np.allclose(o.perturb_normal_det(a, 1.0, n, [-1.0, 0, 1.0], t0), n)
    # the fifth halving would face the camera, but there is no fifth try
    # axis parallel to n: kept
    assert np.allclose(o.perturb_normal_det([0, 0, 1.0], 1.0, n, [0, 0, 1.0], 1.0), n)


# ---------------------------------------------------------------- run start and read-out

def test_topk_mean(o=oracle):
    """Mean of the min(K, n) smallest costs (patchmatch.cu:327-345); 2 without sources."""
    assert o.topk_mean([0.5, 0.1, 0.3, 0.2], 3) == pytest.approx(0.2, rel=1e-6)
    assert o.topk_mean([0.5, 0.1, 0.3, 0.2], 2) == pytest.approx(0.15, rel=1e-6)
    assert o.topk_mean([0.5, 0.1, 0.3, 0.2], 1) == pytest.approx(0.1, rel=1e-6)
    assert o.topk_mean([0.1, 0.3, 0.2], 3) == pytest.approx(0.2, rel=1e-6)  # order of arrival
    assert o.topk_mean([0.3, 0.2, 0.1, 0.05], 3) == pytest.approx(0.35 / 3, rel=1e-6)
    assert o.topk_mean([0.1, 0.1, 0.1, 0.05], 3) == pytest.approx(0.25 / 3, rel=1e-6)  # ties
    assert o.topk_mean([0.8, 0.4], 3) == pytest.approx(0.6, rel=1e-6)  # K > n: mean of both
    assert o.topk_mean([0.8], 3) == pytest.approx(0.8, rel=1e-6)
    assert o.topk_mean([], 3) == 2.0


def test_readout_source_tests(o=oracle):
    """A source counts when the triangulation angle at X is at least the minimum, the surface faces
    the source, the homography exists, the visibility probability E/(E + 0.5) with E = A exp(-m^2 /
    (2 sigma^2)) is at least q_min, and the geometric error is at most the maximum
    (patchmatch.cu:714-733). Reference camera at the origin, X = (0, 0, 10)."""
    X = [0.0, 0.0, 10.0]
    nrm = [0.0, 0.0, -1.0]
    # E = 0.5 for any m (A = 0.5, 1/(2 sigma^2) = 0): visibility exactly 0.5
    k = o.ReadoutParams(math.cos(math.radians(3.0)), 0.5, 0.0, 0.5, 1.0)
    a = 10.0 * math.tan(math.radians(5.0))  # 5 degrees of triangulation
    assert o.readout_source(k, X, [a, 0, 0], nrm, True, 0.0, 1.0)     # visibility = q_min, geo = max
    assert not o.readout_source(k, X, [a, 0, 0], nrm, True, 0.0, 1.0001)
    assert not o.readout_source(k, X, [a, 0, 0], nrm, False, 0.0, 0.0)
    k_hi = o.ReadoutParams(k.cos_filter_tri, 0.5, 0.0, 0.50001, 1.0)
    assert not o.readout_source(k_hi, X, [a, 0, 0], nrm, True, 0.0, 0.0)
    b = 10.0 * math.tan(math.radians(2.0))  # 2 degrees: too small
    assert not o.readout_source(k, X, [b, 0, 0], nrm, True, 0.0, 0.0)
    # a source centre on the ray through X: angle 0 = minimum 0 counts (cos 1 <= cos 0)
    k0 = o.ReadoutParams(1.0, 0.5, 0.0, 0.5, 1.0)
    assert o.readout_source(k0, X, [0, 0, 0], nrm, True, 0.0, 0.0)
    assert not o.readout_source(k0, X, [0, 0, 10.0], nrm, True, 0.0, 0.0)  # centre at X: no direction
    # incidence: the normal perpendicular to X -> c is not facing (n . (c - X) = 0)
    assert not o.readout_source(k0, X, [0, 0, 0], [1.0, 0, 0], True, 0.0, 0.0)
    assert not o.readout_source(k0, X, [0, 0, 0], [0, 0, 1.0], True, 0.0, 0.0)  # facing away
    # visibility with a real emission: sigma = 0.6, A = 1.330950 (math.rs:165), q_min = 0.5:
    # E(m)/(E(m)+0.5) >= 0.5 <=> E(m) >= 0.5 <=> m <= sigma sqrt(2 ln(A / 0.5))
    sig, A = 0.6, 1.330950
    ks = o.ReadoutParams(1.0, A, 1.0 / (2 * sig * sig), 0.5, 1.0)
    m_max = sig * math.sqrt(2 * math.log(A / 0.5))
    assert o.readout_source(ks, X, [0, 0, 0], nrm, True, m_max * 0.999, 0.0)
    assert not o.readout_source(ks, X, [0, 0, 0], nrm, True, m_max * 1.001, 0.0)


def test_count_threshold(o=oracle):
    """Depth kept when count >= max(1, min(min_num_consistent, nsrc)) (densify.rs:446-454)."""
    assert o.count_need(3, 6) == 3
    assert o.count_need(3, 2) == 2
    assert o.count_need(0, 5) == 1
    assert o.count_need(3, 0) == 1
    assert o.count_need(300, 400) == 255
