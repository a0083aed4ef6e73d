"""Statistical pin of the oracle's multi-scale schedule (compute_depth_maps,
crates/dense/src/densify.rs:350-454): on the reference's ground-truth scene (SynthConfig::default(),
3 x 3 cameras 384 x 288, 3 scales, fast profile) the bars of crates/cuda/tests/patchmatch_gpu.rs:96-98
hold: depth median relative error < 0.005, normal median < 3 deg, valid fraction > 0.5.
CPU only (about two minutes on 8 cores)."""
import os

import numpy as np

from cumulus3d_amd import abi, scenes


def test_oracle_meets_reference_ground_truth_bars():
    import oracle

    sc = scenes.make_synth_scene()
    src = oracle.select_sources(sc)
    ranges = oracle.depth_ranges(sc)
    h, w = sc.views[0].gray.shape
    nl = oracle.lib().pmo_num_levels(min(w, h), 3, 64)
    assert nl == 3
    levels = [[abi.LevelData(g, k) for g, k in oracle.build_pyramid(v.gray, v.k, nl)] for v in sc.views]
    pmp, fl = abi.fast_profile()
    inp = scenes.kernel_input(sc, src, ranges, pmp, fl, levels=levels)
    views = list(range(len(sc.views)))
    outs, _, _ = oracle.Oracle(inp, threads=os.cpu_count()).densify(views, abi.Schedule.fast())
    de, ne, total = [], [], 0
    for i, v in enumerate(views):
        d = outs[i].depth.reshape(-1).astype(np.float64)
        t = sc.gt_depth[v].reshape(-1).astype(np.float64)
        total += d.size
        m = (d > 0) & (t > 0)
        de.append(np.abs(d[m] - t[m]) / t[m])
        c = (outs[i].normal.reshape(-1, 3)[m].astype(np.float64) * sc.gt_normal[v].reshape(-1, 3)[m]).sum(1)
        ne.append(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))
    de, ne = np.concatenate(de), np.concatenate(ne)
    valid, dmed, nmed = de.size / total, float(np.median(de)), float(np.median(ne))
    print(f"valid {valid:.3f} depth median {dmed:.2e} normal median {nmed:.2f} deg")
    assert dmed < 0.005
    assert nmed < 3.0
    assert valid > 0.5
