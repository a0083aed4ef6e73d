"""Schedule autotuner of the learned-initialiser path (no reference counterpart).

Picks the cheapest refinement schedule after the learned initialisation that still meets the
accuracy gate against plain PatchStereo on a calibration subset of the views: depth within 1 %
relative on at least `gate` of the plain method's valid pixels. Every candidate keeps a
geometric-consistency round, as the zone workload's method (learned init + PatchStereo +
geometric consistency) does. The same rule picks the cheapest schedule for the control arm that
starts from the reference's random initialisation (patchmatch.cu:668-677), so the two methods are
compared at the same gate. The tile shape of the half step is tuned separately by the engine
(c3d_pm_autotune).
"""
from __future__ import annotations

import dataclasses
from typing import List, Sequence, Tuple

import numpy as np

from . import abi, init_net, pm

# (refine iterations, geometric rounds, geometric iterations, confidence below which the random
# initialisation is used)
CANDIDATES: List[Tuple[int, int, int, float]] = [(0, 1, 1, 0.5), (0, 1, 1, 0.25), (0, 1, 1, 0.0), (0, 1, 2, 0.25), (0, 1, 2, 0.0),
                                                 (1, 1, 1, 0.5), (1, 1, 1, 0.25), (1, 1, 1, 0.0), (1, 1, 2, 0.5), (1, 1, 2, 0.25),
                                                 (1, 1, 2, 0.0), (2, 1, 1, 0.5), (2, 1, 1, 0.25), (2, 1, 2, 0.5), (3, 1, 2, 0.5),
                                                 (4, 1, 2, 0.5), (5, 1, 2, 0.5)]
# (photometric iterations from random initialisation, geometric rounds, geometric iterations)
RANDOM_CANDIDATES: List[Tuple[int, int, int]] = [(0, 1, 1), (0, 1, 2), (1, 1, 1), (1, 1, 2), (2, 1, 1), (2, 1, 2), (3, 1, 1),
                                                 (3, 1, 2), (4, 1, 2), (5, 1, 2), (7, 1, 2)]


def depth_agreement(plain_outs, plain_raw, raw) -> float:
    """Fraction of the plain method's valid pixels (read-out depth > 0) where `raw` is within 1 %
    of the plain raw depth."""
    ok = tot = 0
    for o, dp, dl in zip(plain_outs, plain_raw, raw):
        m = o.depth > 0
        tot += int(m.sum())
        ok += int((np.abs(dl[m] - dp[m]) <= 0.01 * dp[m]).sum())
    return ok / max(tot, 1)


def _choose(rows, gate):
    """rows = [(candidate, agreement, ms)]: the fastest candidate reaching `gate`, else the one with
    the highest agreement (ties: the earlier one)."""
    passing = [r for r in rows if r[1] >= gate]
    if passing:
        return min(passing, key=lambda r: r[2])[0]
    return max(rows, key=lambda r: r[1])[0]


def plain_reference(session: pm.Session, views: Sequence[int], plain: abi.Schedule):
    """The plain method's read-out (before any post-processing) that the candidates are held to."""
    p_outs, _, p_raw, _ = session.densify(views, dataclasses.replace(plain, post=None))
    return p_outs, p_raw


def tune_learned(session: pm.Session, net: init_net.InitNet, views: Sequence[int], plain: abi.Schedule, gate: float = 0.99,
                 candidates: Sequence[Tuple[int, int, int, float]] = CANDIDATES, reference=None):
    """Returns (chosen LearnedSchedule, report rows [(candidate, agreement, device ms)]). Every
    candidate is run and timed; the fastest reaching `gate` wins, else the most accurate."""
    p_outs, p_raw = reference if reference is not None else plain_reference(session, views, plain)
    rows = []
    for (ri, gr, gi, cm) in candidates:
        sch = init_net.LearnedSchedule(refine_iters=ri, geometric_rounds=gr, geometric_iters=gi, conf_min=cm)
        outs, _, raw, tm, _, _ = init_net.densify_learned(session, net, views, sch)
        rows.append(((ri, gr, gi, cm), depth_agreement(p_outs, p_raw, raw), tm.total_ms))
    ri, gr, gi, cm = _choose(rows, gate)
    return init_net.LearnedSchedule(refine_iters=ri, geometric_rounds=gr, geometric_iters=gi, conf_min=cm), rows


def tune_random(session: pm.Session, views: Sequence[int], plain: abi.Schedule, gate: float = 0.99,
                candidates: Sequence[Tuple[int, int, int]] = RANDOM_CANDIDATES, reference=None):
    """The control arm: the fastest schedule from the reference's random initialisation (a plain
    run of fewer iterations) that reaches `gate`. Returns (abi.Schedule, rows)."""
    p_outs, p_raw = reference if reference is not None else plain_reference(session, views, plain)
    rows = []
    for (pi, gr, gi) in candidates:
        sch = abi.Schedule(photometric_iters=pi, geometric_rounds=gr, geometric_iters=gi)
        outs, _, raw, tm = session.densify(views, sch)
        rows.append(((pi, gr, gi), depth_agreement(p_outs, p_raw, raw), tm.total_ms))
    pi, gr, gi = _choose(rows, gate)
    return abi.Schedule(photometric_iters=pi, geometric_rounds=gr, geometric_iters=gi), rows
