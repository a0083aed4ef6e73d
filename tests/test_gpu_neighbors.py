"""GPU scene preparation (c3d_pair_stats, c3d_depth_ranges; PairStats::new and depth_ranges,
crates/dense/src/neighbors.rs:47-189) against the oracle (oracle/scene_oracle.c), through the C ABI.

Bars: shared counts and depth ranges identical; 75th-percentile angles identical up to the last
bit of the float angle (acos is the one library function involved: the device's and the host's
may differ in the last double bit, which survives the rounding to float about once in 2^29), and
identical source lists; on the reference's own scene (host_parts.rs:61-89) the reference's orders.
"""
import math

import numpy as np
import pytest

import oracle
from cumulus3d_amd import neighbors, pm, scenes
from test_scene_oracle import random_tracks, reference_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be():
    b = pm.HipPatchMatch()
    yield b
    b.close()


def test_reference_scene(be):
    sc = reference_scene()
    count, q75 = neighbors.pair_stats(be, sc)
    a1 = math.radians(1.0)
    assert neighbors.select(count, q75, 0, 10, a1) == [1, 3]
    assert neighbors.select(count, q75, 0, 1, a1) == [1]
    assert neighbors.select(count, q75, 0, 10, 0.0) == [2, 1, 3]
    assert neighbors.select(count, q75, 1, 10, a1) == [0, 3]
    r = neighbors.depth_ranges(be, sc)
    assert abs(r[0][0] - 7.5) < 1e-9 and abs(r[0][1] - 12.5) < 1e-9


@pytest.mark.parametrize("nv,npts,seed", [(6, 400, 1), (13, 2000, 2), (3, 5, 3), (108, 100000, 4), (2, 0, 5)])
def test_matches_oracle(be, nv, npts, seed):
    sc = random_tracks(nv, npts, seed)
    s = neighbors.SparseScene(sc)
    count, q75 = neighbors.pair_stats(be, s)
    rc, rq = oracle.pair_stats(s)
    assert np.array_equal(count, rc)
    exact = (q75 == rq).mean()
    print(f"{nv} views, {npts} points: {int(count.sum()) // 2} pairs, q75 bit-identical on {exact:.6f}")
    assert np.allclose(q75, rq, rtol=2.4e-7, atol=0)
    assert neighbors.depth_ranges(be, s) == oracle.depth_ranges(s)
    assert neighbors.select_all(count, q75) == oracle.select_sources(s)


def test_synthetic_scene_sources(be):
    """The reference's accuracy fixture (synthetic.rs): device and oracle agree on every view's
    sources (NeighborParams defaults) and depth range."""
    sc = scenes.make_synth_scene()
    src, ranges = neighbors.sources_and_ranges(be, sc)
    assert src == oracle.select_sources(sc)
    assert ranges == oracle.depth_ranges(sc)
    assert all(len(s) > 0 for s in src) and all(r is not None for r in ranges)
