"""The oracle against the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py): RNG streams, per-hypothesis FP32 and FP64 costs, a random-init
run with a geometric iteration and its read-out counts — bit for bit (FP64 to 1e-12)."""
import ctypes as C
import os
import sys

import numpy as np

import oracle
from cumulus3d_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
from make_golden import ncc_hypotheses, terrain_input  # noqa: E402


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_rng_streams():
    g = _load("rng.npz")
    L = oracle.lib()
    for a, s, d in zip(g["args"], g["states"], g["draws"]):
        st = C.c_uint64(L.pmo_rng_state(*[int(q) for q in a]))
        assert st.value == int(s)
        assert np.array_equal(np.array([L.pmo_rng_next(C.byref(st)) for _ in range(8)], np.float32), d)


def test_costs():
    g = _load("ncc_96x54.npz")
    sc, inp = terrain_input(96, 54, seed=21)
    xy, planes = ncc_hypotheses(sc)
    assert np.array_equal(xy, g["xy"]) and np.array_equal(planes.view(np.uint32), g["planes"].view(np.uint32))
    o = oracle.Oracle(inp, threads=4)
    f32 = np.array([o.ncc_cost_f32(0, 0, 0, int(x), int(y), p) for (x, y), p in zip(xy, planes)], np.float32)
    f64 = np.array([o.ncc_cost_f64(0, 0, 0, int(x), int(y), p.astype(np.float64)) for (x, y), p in zip(xy, planes)])
    assert np.array_equal(f32.view(np.uint32), g["cost_f32"].view(np.uint32))
    assert np.allclose(f64, g["cost_f64"], rtol=0, atol=1e-12)
    ok = (g["cost_f64"] < 2) & (g["cost_f32"] < 2)
    assert ok.sum() > 100 and np.abs(g["cost_f32"][ok] - g["cost_f64"][ok]).max() < 1e-4


def test_run_and_readout():
    g = _load("run_64x48.npz")
    sc, inp = terrain_input(64, 48, seed=22)
    o = oracle.Oracle(inp, threads=4)
    st = [abi.ViewState(64, 48) for _ in range(3)]
    o.run([0, 1, 2], st, abi.run_params(0, False, True, 3, 1, False))
    snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in st])
    o.run([0, 1, 2], st, abi.run_params(0, True, False, 1, 2, False), snap)
    cnt = o.filter([0, 1, 2], st, snap)
    for v in range(3):
        for key, arr in (("depth", st[v].depth), ("normal", st[v].normal), ("cost", st[v].cost)):
            assert np.array_equal(arr.view(np.uint32), g[key][v].view(np.uint32)), (v, key)
        assert np.array_equal(cnt[v], g["count"][v])
