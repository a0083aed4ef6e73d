"""The device against the committed golden fixtures (tests/golden/): per-hypothesis costs and a
random-init run with a geometric iteration and its read-out, bit for bit, through the C ABI."""
import os
import sys

import numpy as np
import pytest

from cumulus3d_amd import abi, pm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
from make_golden import ncc_hypotheses, terrain_input  # noqa: E402

pytestmark = pytest.mark.gpu


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_costs_match_golden():
    g = _load("ncc_96x54.npz")
    sc, inp = terrain_input(96, 54, seed=21)
    xy, planes = ncc_hypotheses(sc)
    be = pm.HipPatchMatch()
    with be.begin(inp) as s:
        got = s.pair_costs(0, 0, 0, xy, planes)
    be.close()
    assert np.array_equal(got.view(np.uint32), g["cost_f32"].view(np.uint32))


def test_run_matches_golden():
    g = _load("run_64x48.npz")
    sc, inp = terrain_input(64, 48, seed=22)
    be = pm.HipPatchMatch()
    st = [abi.ViewState(64, 48) for _ in range(3)]
    with be.begin(inp) as s:
        s.run([0, 1, 2], st, abi.run_params(0, False, True, 3, 1, False))
        snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in st])
        s.run([0, 1, 2], st, abi.run_params(0, True, False, 1, 2, False), snap)
        cnt = s.filter([0, 1, 2], st, snap)
    be.close()
    for v in range(3):
        for key, arr in (("depth", st[v].depth), ("normal", st[v].normal), ("cost", st[v].cost)):
            assert np.array_equal(arr.view(np.uint32), g[key][v].view(np.uint32)), (v, key)
        assert np.array_equal(cnt[v], g["count"][v])
