"""GPU agreement fusion (c3d_fuse_consistency) against the oracle (oracle/fusion_oracle.c): the
same points bit for bit (positions, normals, colours, keys) in the same order, on exact plane depth
maps and on depth maps the engine itself produced for a terrain formation."""
import numpy as np
import pytest

from cumulus3d_amd import abi, fusion, pm, scenes

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert len(a.positions) == len(b.positions)
    assert np.array_equal(a.keys, b.keys)
    assert np.array_equal(a.positions.view(np.uint32), b.positions.view(np.uint32))
    assert np.array_equal(a.normals.view(np.uint32), b.normals.view(np.uint32))
    assert np.array_equal(a.colors, b.colors)


def test_plane_fusion_bit_exact():
    import oracle
    import test_fusion_oracle as t

    sc, depth, normal = t._plane_scene()
    src = scenes.nearest_sources(sc, [4] * len(sc.views))
    keep = []
    views = fusion.fusion_views(sc, depth, normal, [np.full_like(d, 0.3) for d in depth], src, src, keep)
    for prm in [fusion.FusionParams(min_consistent_views=2), fusion.FusionParams(min_consistent_views=2, mark_radius=1)]:
        g = fusion.fuse(pm.HipPatchMatch(), views, prm, 10 ** 6)
        r = oracle.fuse_consistency(views, prm, 10 ** 6)
        assert len(g.positions) > 1000
        _same(g, r)


def test_terrain_fusion_bit_exact():
    import oracle

    rng = np.random.default_rng(5)
    cams = scenes.formation_rig(4, 3, spacing=6.0)
    sc = scenes.make_mixed_scene(cams, [(192, 108)] * len(cams), seed=9)
    src = scenes.nearest_sources(sc, [int(k) for k in rng.integers(3, 6, len(cams))])
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, [(15.0, 120.0)] * len(cams), pmp, fl)
    be = pm.HipPatchMatch()
    with be.begin(inp) as s:
        outs, _, _, _ = s.densify(list(range(len(cams))), abi.Schedule(photometric_iters=5, geometric_rounds=1, geometric_iters=2))
    keep = []
    views = fusion.fusion_views(sc, [o.depth for o in outs], [o.normal for o in outs], [o.cost for o in outs], src, src, keep)
    prm = fusion.FusionParams(min_consistent_views=2)
    g = fusion.fuse(be, views, prm, 10 ** 6)
    r = oracle.fuse_consistency(views, prm, 10 ** 6)
    print(f"terrain fusion: {len(g.positions)} points")
    assert len(g.positions) > 500
    _same(g, r)
    # a too small capacity is reported, not overrun
    with pytest.raises(pm.PmError) as e:
        fusion.fuse(be, views, prm, 10)
    assert e.value.status == abi.C3D_INVALID_ARGUMENT


def test_fusion_without_valid_depth():
    """Every depth invalid (zero, negative or NaN): no point, on the device as in the oracle."""
    import oracle
    import test_fusion_oracle as t

    sc, depth, normal = t._plane_scene()
    src = scenes.nearest_sources(sc, [4] * len(sc.views))
    bad = []
    for i, d in enumerate(depth):
        z = np.zeros_like(d)
        z[::3] = -1.0
        z[1::3] = np.nan
        bad.append(z)
    keep = []
    views = fusion.fusion_views(sc, bad, normal, [np.full_like(d, 0.3) for d in depth], src, src, keep)
    prm = fusion.FusionParams(min_consistent_views=2)
    g = fusion.fuse(pm.HipPatchMatch(), views, prm, 10 ** 5)
    r = oracle.fuse_consistency(views, prm, 10 ** 5)
    assert len(r.positions) == 0
    _same(g, r)
