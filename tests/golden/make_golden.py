"""Writes the golden fixtures of tests/golden/ from the CPU oracle (oracle/): small seeded inputs
and the oracle's outputs, so that a later change of the oracle or of the device code shows up as
a difference against data fixed today. The inputs are regenerated from their seeds by the tests
(the terrain renderer is deterministic); only outputs and the hypothesis lists are stored.

  python tests/golden/make_golden.py     (rewrites tests/golden/*.npz)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cumulus3d_amd import abi, scenes  # noqa: E402

SRC3 = [[1, 2], [0, 2], [0, 1]]


def terrain_input(width, height, seed):
    """A 3-view oblique terrain scene and its kernel input (fast profile, depth 20-80 m)."""
    rig = scenes.oblique_rig(3, baseline=6.0)
    sc = scenes.make_terrain_scene(rig, width, height, focal=500.0 * width / 640.0, seed=seed)
    pmp, fl = abi.fast_profile()
    return sc, scenes.kernel_input(sc, SRC3, [(20.0, 80.0)] * 3, pmp, fl)


def ncc_hypotheses(sc, n=200, seed=7):
    """Hypotheses at random interior pixels of view 0: the true plane, perturbed in depth."""
    rng = np.random.default_rng(seed)
    h, w = sc.views[0].gray.shape
    xy = np.stack([rng.integers(5, w - 5, n), rng.integers(5, h - 5, n)], 1).astype(np.int32)
    planes = np.array([[*sc.gt_normal[0][y, x], sc.gt_depth[0][y, x] * (1.0 + 0.02 * rng.standard_normal())]
                       for x, y in xy], np.float32)
    return xy, planes


def main():
    # RNG streams (kernel.rs:248-259)
    L = oracle.lib()
    import ctypes as C

    args = [(0x5EED0FDE75, 0x1234, 0, 1, 0, 0, 17), (42, 7, 2, 3, 5, 1, 123456), (0, 0, 0, 0, 0xFFFF, 0, 0)]
    states, draws = [], []
    for a in args:
        st = C.c_uint64(L.pmo_rng_state(*a))
        states.append(st.value)
        draws.append([L.pmo_rng_next(C.byref(st)) for _ in range(8)])
    np.savez(os.path.join(HERE, "rng.npz"), args=np.array(args, np.uint64), states=np.array(states, np.uint64),
             draws=np.array(draws, np.float32))

    # per-hypothesis costs on a 96 x 54 scene: FP32 (the device's arithmetic) and FP64
    sc, inp = terrain_input(96, 54, seed=21)
    o = oracle.Oracle(inp, threads=4)
    xy, planes = ncc_hypotheses(sc)
    f32 = np.array([o.ncc_cost_f32(0, 0, 0, int(x), int(y), p) for (x, y), p in zip(xy, planes)], np.float32)
    f64 = np.array([o.ncc_cost_f64(0, 0, 0, int(x), int(y), p.astype(np.float64)) for (x, y), p in zip(xy, planes)])
    np.savez(os.path.join(HERE, "ncc_96x54.npz"), xy=xy, planes=planes, cost_f32=f32, cost_f64=f64)

    # a random-init run of 3 iterations and a geometric iteration with read-out on 64 x 48 views
    sc, inp = terrain_input(64, 48, seed=22)
    o = oracle.Oracle(inp, threads=4)
    st = [abi.ViewState(64, 48) for _ in range(3)]
    o.run([0, 1, 2], st, abi.run_params(0, False, True, 3, 1, False))
    snap = abi.DepthSnapshot(3, 0, [x.depth.copy() for x in st])
    o.run([0, 1, 2], st, abi.run_params(0, True, False, 1, 2, False), snap)
    cnt = o.filter([0, 1, 2], st, snap)
    np.savez(os.path.join(HERE, "run_64x48.npz"), depth=np.stack([x.depth for x in st]), normal=np.stack([x.normal for x in st]),
             cost=np.stack([x.cost for x in st]), count=np.stack(cnt))
    print("wrote rng.npz, ncc_96x54.npz, run_64x48.npz")


if __name__ == "__main__":
    main()
