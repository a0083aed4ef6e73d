"""The learned-schedule autotuner's choice rule (cumulus3d_amd/autotune.py), on the CPU with a
stand-in session: candidates are tried in cost order, the first whose depth agreement with plain
PatchStereo reaches the gate wins, otherwise the best one; the plain reference is computed
without post-processing."""
import numpy as np

from cumulus3d_amd import abi, autotune, init_net


class FakeOut:
    def __init__(self, d):
        self.depth = d


class FakeSession:
    """Plain depth 10 everywhere; a learned schedule's raw depth is off by 2 % on a fraction of
    the pixels that shrinks with its refinement iterations."""

    def __init__(self):
        self.plain_calls = []

    def densify(self, views, sched, download=True):
        self.plain_calls.append(sched.post)
        d = np.full((10, 10), 10.0, np.float32)
        return [FakeOut(d) for _ in views], None, [d.copy() for _ in views], None


def test_first_candidate_meeting_the_gate_wins(monkeypatch):
    bad = {1: 30, 2: 1, 3: 0}  # pixels of 100 off by 2 %, per refinement iterations

    def fake_learned(session, net, views, sch, download=True, want_init=False):
        raw = []
        for _ in views:
            d = np.full((10, 10), 10.0, np.float32)
            d.reshape(-1)[: bad.get(sch.refine_iters, 0)] = 10.2
            raw.append(d)
        tm = abi.c3d_timing()
        tm.total_ms = 10.0 * sch.refine_iters
        return None, None, raw, tm, None, None

    monkeypatch.setattr(init_net, "densify_learned", fake_learned)
    s = FakeSession()
    plain = abi.Schedule(post=abi.PostParams.fast())
    cands = [(1, 0, 0, 0.5), (2, 0, 0, 0.5), (3, 0, 0, 0.5)]
    sch, rows = autotune.tune_learned(s, None, [0, 1], plain, gate=0.99, candidates=cands)
    assert (sch.refine_iters, sch.geometric_rounds) == (2, 0)
    assert [r[0] for r in rows] == cands[:2]
    assert abs(rows[0][1] - 0.70) < 1e-9 and abs(rows[1][1] - 0.99) < 1e-9
    assert s.plain_calls == [None]  # the plain reference is taken before post-processing
    # nothing reaches the gate: the best candidate
    sch, rows = autotune.tune_learned(s, None, [0], plain, gate=1.01, candidates=cands)
    assert sch.refine_iters == 3 and len(rows) == 3
