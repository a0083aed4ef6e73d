"""Schedule autotuner of the learned-initialiser path (no reference counterpart).

Picks the cheapest refinement schedule after the learned initialisation that still meets the
accuracy gate against plain PatchStereo on a calibration subset of the views: depth within 1 %
relative on at least `gate` of the plain method's valid pixels. The tile height of the half step
is tuned separately by the engine (c3d_pm_autotune).
"""
from __future__ import annotations

import dataclasses
import time
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi, init_net, pm

# (refine iterations, geometric rounds, geometric iterations, confidence below which the random
# initialisation is used), cheapest first
CANDIDATES: List[Tuple[int, int, int, float]] = [(1, 0, 0, 0.5), (2, 0, 0, 0.5), (1, 1, 1, 0.5), (1, 1, 1, 0.25), (1, 1, 1, 0.0),
                                                 (2, 1, 1, 0.5), (2, 1, 1, 0.25), (2, 1, 2, 0.5), (3, 1, 2, 0.5), (4, 1, 2, 0.5),
                                                 (5, 1, 2, 0.5)]


def depth_agreement(plain_outs, plain_raw, raw) -> float:
    ok = tot = 0
    for o, dp, dl in zip(plain_outs, plain_raw, raw):
        m = o.depth > 0
        tot += int(m.sum())
        ok += int((np.abs(dl[m] - dp[m]) <= 0.01 * dp[m]).sum())
    return ok / max(tot, 1)


def tune_learned(session: pm.Session, net: init_net.InitNet, views: Sequence[int], plain: abi.Schedule, gate: float = 0.99,
                 candidates: Sequence[Tuple[int, int, int, float]] = CANDIDATES):
    """Returns (chosen LearnedSchedule, report rows [(candidate, agreement, ms)]). The first
    candidate (in cost order) whose agreement reaches `gate` wins; if none does, the one with the
    highest agreement."""
    # the plain method's read-out mask before any post-processing
    p_outs, _, p_raw, _ = session.densify(views, dataclasses.replace(plain, post=None))
    rows = []
    best = None
    for (ri, gr, gi, cm) in candidates:
        sch = init_net.LearnedSchedule(refine_iters=ri, geometric_rounds=gr, geometric_iters=gi, conf_min=cm)
        t = time.perf_counter()
        outs, _, raw, tm, _, _ = init_net.densify_learned(session, net, views, sch)
        ms = (time.perf_counter() - t) * 1e3
        agree = depth_agreement(p_outs, p_raw, raw)
        rows.append(((ri, gr, gi, cm), agree, tm.total_ms))
        if agree >= gate:
            return sch, rows
        if best is None or agree > best[1]:
            best = (sch, agree)
    return best[0], rows
