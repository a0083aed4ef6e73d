"""Result downloads through the page-locked staging buffers (two buffers that the device and the
host threads fill and empty in turn). With the buffers shrunk to a few KB ($C3D_PM_STAGE_BYTES) a
small job switches between them dozens of times and splits every array across buffers; the
results equal those downloaded with the default 256 MB buffers (one switch) and the oracle's."""
import os
import sys

import numpy as np
import pytest

from cumulus3d_amd import abi, pm, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _input(width=192, height=108, seed=4):
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 640.0, seed=seed)
    src = [[j for j in range(3) if j != v] for v in range(3)]
    pmp, fl = abi.fast_profile()
    pmp = abi.PmParams(**{**pmp.__dict__, "window_radius": 5, "window_step": 2})
    return sc, scenes.kernel_input(sc, src, [(15.0, 120.0)] * 3, pmp, fl)


def _bits(a):
    return np.asarray(a).view(np.uint32)


def _densify_and_run(inp):
    """A densify call (states, counts, raw depth) and a trait-level run with states in and out."""
    sched = abi.Schedule(photometric_iters=3, geometric_rounds=1, geometric_iters=2)
    be = pm.HipPatchMatch()
    with be.begin(inp) as s:
        outs, counts, raws, _ = s.densify([0, 1, 2], sched)
        st = [o.copy() for o in outs]
        s.run([0, 1, 2], st, abi.run_params(0, False, False, 1, 7, False))
    be.close()
    return outs, counts, raws, st


def test_downloads_through_many_small_staging_buffers_equal_the_default_and_the_oracle(monkeypatch):
    import oracle

    sc, inp = _input()
    monkeypatch.delenv("C3D_PM_STAGE_BYTES", raising=False)
    ref = _densify_and_run(inp)
    # 4 KB + 4 bytes: never a multiple of an array, so pieces straddle the buffers
    monkeypatch.setenv("C3D_PM_STAGE_BYTES", "4100")
    got = _densify_and_run(inp)
    monkeypatch.delenv("C3D_PM_STAGE_BYTES")
    (o1, c1, r1, s1), (o2, c2, r2, s2) = ref, got
    for v in range(3):
        for x, y in ((o1[v].depth, o2[v].depth), (o1[v].normal, o2[v].normal), (o1[v].cost, o2[v].cost), (r1[v], r2[v]),
                     (s1[v].depth, s2[v].depth), (s1[v].normal, s2[v].normal), (s1[v].cost, s2[v].cost)):
            assert np.array_equal(_bits(x), _bits(y))
        assert np.array_equal(c1[v], c2[v])
    o = oracle.Oracle(inp, threads=16)
    ro, rc, rraw = o.densify([0, 1, 2], abi.Schedule(photometric_iters=3, geometric_rounds=1, geometric_iters=2))
    outs, counts, raws, _ = got
    for v in range(3):
        assert np.array_equal(_bits(outs[v].depth), _bits(ro[v].depth))
        assert np.array_equal(_bits(outs[v].normal), _bits(ro[v].normal))
        assert np.array_equal(_bits(raws[v]), _bits(rraw[v]))
        assert np.array_equal(counts[v], rc[v])
