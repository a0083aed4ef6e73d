"""GPU read-out post-processing (c3d_remove_speckles, c3d_fill_holes and the post stage of
c3d_pm_densify; postproc.rs:77-204, densify.rs:455-462) against the oracle (oracle/post_oracle.c),
through the C ABI.

Bars: bit-exact depth, normal and cost and identical counts, on seeded maps from 1 x 200 up to
1920 x 1080 (speckles with tied costs, NaN and negative depths, holes by gray edges), and on the
full multi-scale schedule with the fast profile's post-processing, eager and captured in a graph.
"""
import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes
from test_post_oracle import speckle_map

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def be():
    b = pm.HipPatchMatch()
    yield b
    b.close()


@pytest.mark.parametrize("h,w,seed,frac", [(30, 41, 1, 2e-2), (1, 200, 4, 3e-2), (50, 1, 5, 1e-1), (256, 384, 6, 2e-3),
                                           (1080, 1920, 8, 2e-4), (64, 64, 9, 1.0 / 4096)])
def test_speckles_bit_exact(be, h, w, seed, frac):
    import oracle

    d, nrm, cost = speckle_map(h, w, seed)
    d[0, 0] = np.nan
    p = abi.PostParams(speckle_min_area_frac=frac)
    dg, ng = d.copy(), nrm.copy()
    dr, nr = d.copy(), nrm.copy()
    got = be.remove_speckles(dg, ng, cost, p)
    want = oracle.remove_speckles(dr, nr, cost, p)
    print(f"{h}x{w}: removed {got} (oracle {want})")
    assert got == want
    assert np.array_equal(_bits(dg), _bits(dr)) and np.array_equal(_bits(ng), _bits(nr))


def _hole_map(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    k = [0.9 * w, 0.9 * w, (w - 1) / 2.0, (h - 1) / 2.0]
    d = (25.0 + 5.0 * np.sin(xx / 37.0) + 3.0 * np.cos(yy / 23.0)).astype(np.float32)
    n = rng.standard_normal((h, w, 3)).astype(np.float32)
    n[..., 2] = -np.abs(n[..., 2]) - 1.5
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    holes = rng.random((h, w)) < 0.25
    for _ in range(max(1, h * w // 2000)):
        y0, x0, r = rng.integers(0, h), rng.integers(0, w), rng.integers(2, 12)
        holes[max(0, y0 - r):y0 + r, max(0, x0 - r):x0 + r] = True
    d[holes] = 0.0
    d[rng.random((h, w)) < 0.01] = np.nan
    cost = rng.uniform(0.0, 1.5, (h, w)).astype(np.float32)
    raw = rng.uniform(0.0, 1.5, (h, w)).astype(np.float32)
    # gray with edges: blocks of constant value plus gentle noise
    gray = ((xx // 17 * 37 + yy // 13 * 53) % 256 + rng.integers(0, 8, (h, w))).clip(0, 255).astype(np.uint8)
    return d, n, cost, raw, gray, k


@pytest.mark.parametrize("h,w,seed,frac", [(40, 60, 1, 0.1), (1, 300, 2, 0.05), (270, 480, 3, 0.01), (1080, 1920, 4, 0.01)])
def test_fill_bit_exact(be, h, w, seed, frac):
    import oracle

    d, n, cost, raw, gray, k = _hole_map(h, w, seed)
    p = abi.PostParams(fill_holes=True, fill_max_frac=frac)
    a = [d.copy(), n.copy(), cost.copy()]
    b = [d.copy(), n.copy(), cost.copy()]
    got = be.fill_holes(a[0], a[1], a[2], raw, gray, k, p)
    want = oracle.fill_holes(b[0], b[1], b[2], raw, gray, k, p)
    print(f"{h}x{w}: filled {got} (oracle {want}) of {int((~(d > 0)).sum())} invalid")
    assert got == want and got > 0
    for x, y in zip(a, b):
        assert np.array_equal(_bits(x), _bits(y))


def _terrain(width, height, nviews=3, seed=7):
    rig = scenes.oblique_rig(nviews, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 640.0, seed=seed)
    src = [[j for j in range(nviews) if j != v] for v in range(nviews)]
    return sc, src


@pytest.mark.parametrize("graph", [0, 1])
def test_densify_post_bit_exact(graph):
    """Three scales of a 320 x 256 terrain with the fast profile's speckle removal and hole filling
    after the threshold, against the oracle's schedule."""
    import oracle

    sc, src = _terrain(320, 256)
    opts = abi.c3d_pm_options()
    opts.use_graph = graph
    be = pm.HipPatchMatch(0, opts)
    nl = pm.num_levels(256, 3, 64)
    levels = [be.build_pyramid(v.gray, v.k, nl) for v in sc.views]
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, [(20.0, 80.0)] * 3, pmp, fl, levels=levels)
    post = abi.PostParams.fast()
    post.speckle_min_area_frac = 2e-3  # components below 164 pixels at 320 x 256
    sched = abi.Schedule(post=post)
    plain = abi.Schedule()
    with be.begin(inp) as s:
        p_out, _, _, _ = s.densify([0, 1, 2], plain)
        go, gc, graw, tm = s.densify([0, 1, 2], sched)
        go2, _, _, tm2 = s.densify([0, 1, 2], sched)
    ro, rc, rraw = oracle.Oracle(inp, threads=16).densify([0, 1, 2], sched)
    print(f"removed {tm.post_removed}, filled {tm.post_filled}, post {tm.post_ms:.2f} ms")
    assert tm.post_removed > 0 and tm.post_filled > 0
    for v in range(3):
        assert np.array_equal(_bits(graw[v]), _bits(rraw[v]))
        assert np.array_equal(gc[v], rc[v])
        for x, y in ((go[v].depth, ro[v].depth), (go[v].normal, ro[v].normal), (go[v].cost, ro[v].cost)):
            assert np.array_equal(_bits(x), _bits(y))
        assert np.array_equal(_bits(go2[v].depth), _bits(go[v].depth))
    # every changed pixel was erased or filled (an erased pixel may be filled again, as in the
    # reference's own test)
    changed = sum(int((_bits(go[v].depth) != _bits(p_out[v].depth)).sum()) for v in range(3))
    assert 0 < changed <= tm.post_removed + tm.post_filled
    be.close()


@pytest.mark.parametrize("kind", ["all_invalid", "one_component"])
def test_degenerate_maps(be, kind):
    """A map with no valid pixel and a map that is one constant component: both filters equal the
    oracle (nothing to erase or fill from in the first, one component larger than any speckle and
    no hole in the second)."""
    import oracle

    h, w = 90, 130
    rng = np.random.default_rng(31)
    d = np.zeros((h, w), np.float32) if kind == "all_invalid" else np.full((h, w), 17.5, np.float32)
    n = np.zeros((h, w, 3), np.float32)
    n[..., 2] = -1.0
    cost = rng.uniform(0.0, 1.5, (h, w)).astype(np.float32)
    raw = rng.uniform(0.0, 1.5, (h, w)).astype(np.float32)
    gray = rng.integers(0, 256, (h, w)).astype(np.uint8)
    k = [0.9 * w, 0.9 * w, (w - 1) / 2.0, (h - 1) / 2.0]
    p = abi.PostParams(fill_holes=True)
    a = [d.copy(), n.copy(), cost.copy()]
    b = [d.copy(), n.copy(), cost.copy()]
    assert be.remove_speckles(a[0], a[1], a[2], p) == oracle.remove_speckles(b[0], b[1], b[2], p) == 0
    assert be.fill_holes(a[0], a[1], a[2], raw, gray, k, p) == oracle.fill_holes(b[0], b[1], b[2], raw, gray, k, p) == 0
    for x, y in zip(a, b):
        assert np.array_equal(_bits(x), _bits(y))
