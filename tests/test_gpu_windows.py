"""GPU parity of the any-radius kernels (window radius 1..20, step 1 or 2; patchmatch.rs:292-294)
against the CPU oracle, through the C ABI.

The windows with a register-resident build ((5, 2), (4, 2), (3, 2), (2, 2), (3, 1), (2, 1),
(1, 1)) also have the any-radius path (c3d_pm_options.any_radius = 1); both must give identical
results. Every other window, the quality profile's (5, 1) among them, runs only the any-radius path
and must equal the oracle bit for bit; its per-hypothesis costs must be within 1e-4 of the FP64
evaluation. The reference's ground-truth bars (patchmatch_gpu.rs:96-98) are checked for the quality
profile (window 5 step 1, 7 / 6 iterations, restorer random init, 2 geometric rounds of 6, median).
"""
import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes

pytestmark = pytest.mark.gpu

SRC3 = [[1, 2], [0, 2], [0, 1]]


def _input(width, height, wr, step, seed=2):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 640.0, seed=seed)
    pmp, fl = abi.fast_profile()
    pmp = abi.PmParams(**{**pmp.__dict__, "window_radius": wr, "window_step": step})
    return sc, scenes.kernel_input(sc, SRC3, [(20.0, 80.0)] * 3, pmp, fl)


def _oracle(inp):
    import oracle

    return oracle.Oracle(inp, threads=16)


def _run(be, inp, iters=3, geo=True):
    views = [0, 1, 2]
    st = [abi.ViewState(*inp.level_shape(v, 0)[::-1]) for v in views]
    with be.begin(inp) as s:
        s.run(views, st, abi.run_params(0, False, True, iters, 1, False))
        if geo:
            snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in st])
            s.run(views, st, abi.run_params(0, True, False, 1, 2, False), snap)
            cnt = s.filter(views, st, snap)
        else:
            cnt = None
    return st, cnt


def _oracle_run(inp, iters=3, geo=True):
    o = _oracle(inp)
    views = [0, 1, 2]
    st = [abi.ViewState(*inp.level_shape(v, 0)[::-1]) for v in views]
    o.run(views, st, abi.run_params(0, False, True, iters, 1, False))
    cnt = None
    if geo:
        snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in st])
        o.run(views, st, abi.run_params(0, True, False, 1, 2, False), snap)
        cnt = o.filter(views, st, snap)
    return st, cnt


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_any_radius_path_equals_register_path():
    """(5, 2) and (3, 1) through both kernel families: identical states and counts."""
    for wr, step in [(5, 2), (3, 1)]:
        sc, inp = _input(192, 108, wr, step)
        a, ca = _run(pm.HipPatchMatch(), inp)
        opts = abi.c3d_pm_options()
        opts.any_radius = 1
        b, cb = _run(pm.HipPatchMatch(0, opts), inp)
        for x, y, u, v in zip(a, b, ca, cb):
            assert _same(x.depth, y.depth) and _same(x.normal, y.normal) and _same(x.cost, y.cost)
            assert np.array_equal(u, v)


@pytest.mark.parametrize("wr,step", [(5, 1), (7, 2), (20, 2), (9, 1)])
def test_any_radius_bit_exact(wr, step):
    w, h = (128, 72) if wr * step < 10 or step == 2 else (96, 64)
    sc, inp = _input(w, h, wr, step, seed=wr)
    iters = 2 if wr >= 9 else 3
    g, cg = _run(pm.HipPatchMatch(), inp, iters=iters, geo=wr <= 7)
    r, cr = _oracle_run(inp, iters=iters, geo=wr <= 7)
    for i in range(3):
        assert _same(g[i].depth, r[i].depth)
        assert _same(g[i].normal, r[i].normal)
        assert _same(g[i].cost, r[i].cost)
        if cg is not None:
            assert np.array_equal(cg[i], cr[i])


def test_quality_window_costs_match_fp64():
    sc, inp = _input(320, 180, 5, 1, seed=9)
    o = _oracle(inp)
    rng = np.random.default_rng(4)
    n = 1500
    xs, ys = rng.integers(0, 320, n), rng.integers(0, 180, n)
    planes = np.zeros((n, 4), np.float32)
    for i, (x, y) in enumerate(zip(xs, ys)):
        nrm = np.asarray(sc.gt_normal[0][y, x], float) + rng.normal(0, 0.05, 3)
        nrm /= np.linalg.norm(nrm)
        planes[i] = [*nrm, float(sc.gt_depth[0][y, x]) * rng.uniform(0.98, 1.02)]
    xy = np.stack([xs, ys], 1).astype(np.int32)
    with pm.HipPatchMatch().begin(inp) as s:
        got = s.pair_costs(0, 0, 0, xy, planes)
    ref = np.array([o.ncc_cost_f64(0, 0, 0, int(x), int(y), p) for (x, y), p in zip(xy, planes.astype(float))])
    both = (ref < 2.0) & (got < 2.0)
    assert both.sum() > 700
    err = np.abs(got[both] - ref[both]).max()
    print(f"(5, 1): {both.sum()} costs, max |gpu - fp64| = {err:.2e}")
    assert err < 1e-4


def test_tile_rows_identical_and_autotune():
    """Half-step workgroups of 1, 2 and 4 rows, and 4 rows with two lanes per pixel (register and
    any-radius windows), give identical states and read-out counts (photometric and geometric
    runs); the autotuner times every shape and keeps one."""
    sc, inp = _input(192, 108, 5, 2)
    outs = []
    for rows, tpp, anyr in ((1, 1, 0), (2, 1, 0), (4, 1, 0), (4, 2, 0), (4, 2, 1), (1, 1, 1)):
        opts = abi.c3d_pm_options()
        opts.tile_rows = rows
        opts.threads_per_pixel = tpp
        opts.any_radius = anyr  # the any-radius kernels with two lanes per pixel too
        outs.append(_run(pm.HipPatchMatch(0, opts), inp))
    for other in outs[1:]:
        for x, y in zip(outs[0][0], other[0]):
            assert _same(x.depth, y.depth) and _same(x.normal, y.normal) and _same(x.cost, y.cost)
        for a, b in zip(outs[0][1], other[1]):
            assert np.array_equal(a, b)
    with pm.HipPatchMatch().begin(inp) as s:
        shape, ms = s.autotune([0, 1, 2])
    print(f"autotune: (rows, lanes per pixel) {shape}, ms {ms}")
    assert shape in abi.TILE_SHAPES and len(ms) == len(abi.TILE_SHAPES) and min(ms) > 0
    bad = abi.c3d_pm_options()
    bad.tile_rows, bad.threads_per_pixel = 2, 2
    with pytest.raises(pm.PmError):
        pm.HipPatchMatch(0, bad)
