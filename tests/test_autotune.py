"""The learned-schedule autotuner's choice rule (cumulus3d_amd/autotune.py), on the CPU with a
stand-in session: every candidate is run and timed, the fastest whose depth agreement with plain
PatchStereo reaches the gate wins, otherwise the most accurate; the plain reference is computed
without post-processing; the random-initialisation control arm is chosen by the same rule."""
import numpy as np

from cumulus3d_amd import abi, autotune, init_net


class FakeOut:
    def __init__(self, d):
        self.depth = d


def _raw(bad):
    d = np.full((10, 10), 10.0, np.float32)
    d.reshape(-1)[:bad] = 10.2  # 2 % off on `bad` of 100 pixels
    return d


class FakeSession:
    """Plain (7 photometric iterations) has depth 10 everywhere; a random-init schedule of fewer
    iterations is off by 2 % on 40 pixels minus 20 per extra iteration."""

    def __init__(self):
        self.plain_calls = []

    def densify(self, views, sched, download=True):
        self.plain_calls.append(sched.post)
        bad = 0 if sched.photometric_iters >= 7 else max(0, 40 - 20 * (sched.photometric_iters - 1) - 5 * sched.geometric_iters)
        tm = abi.c3d_timing()
        tm.total_ms = 10.0 * sched.photometric_iters + 15.0 * sched.geometric_rounds * sched.geometric_iters
        d = np.full((10, 10), 10.0, np.float32)
        return [FakeOut(d) for _ in views], None, [_raw(bad) for _ in views], tm


def test_fastest_candidate_meeting_the_gate_wins(monkeypatch):
    # pixels of 100 off by 2 %, and device ms, per (refine, geometric iterations, confidence)
    table = {(1, 1, 0.5): (3, 25.0), (1, 1, 0.25): (1, 26.0), (1, 2, 0.5): (0, 40.0), (2, 1, 0.5): (0, 35.0)}

    def fake_learned(session, net, views, sch, download=True, want_init=False):
        bad, ms = table[(sch.refine_iters, sch.geometric_iters, sch.conf_min)]
        tm = abi.c3d_timing()
        tm.total_ms = ms
        return None, None, [_raw(bad) for _ in views], tm, None, None

    monkeypatch.setattr(init_net, "densify_learned", fake_learned)
    s = FakeSession()
    plain = abi.Schedule(photometric_iters=7, geometric_rounds=1, geometric_iters=2, post=abi.PostParams.fast())
    cands = [(1, 1, 1, 0.5), (1, 1, 1, 0.25), (1, 1, 2, 0.5), (2, 1, 1, 0.5)]
    sch, rows = autotune.tune_learned(s, None, [0, 1], plain, gate=0.99, candidates=cands)
    # (1,1,1,0.5) is fastest but fails (0.97); of the three passing, (1,1,1,0.25) is fastest
    assert (sch.refine_iters, sch.geometric_rounds, sch.geometric_iters, sch.conf_min) == (1, 1, 1, 0.25)
    assert [r[0] for r in rows] == cands  # every candidate is measured
    assert abs(rows[0][1] - 0.97) < 1e-9 and abs(rows[1][1] - 0.99) < 1e-9 and rows[2][1] == 1.0
    assert [r[2] for r in rows] == [25.0, 26.0, 40.0, 35.0]
    assert s.plain_calls == [None]  # the plain reference is taken before post-processing
    # nothing reaches the gate: the most accurate candidate, the earlier one on a tie
    sch, rows = autotune.tune_learned(s, None, [0], plain, gate=1.01, candidates=cands)
    assert (sch.refine_iters, sch.geometric_iters) == (1, 2)


def test_random_control_arm_uses_the_same_rule():
    s = FakeSession()
    plain = abi.Schedule(photometric_iters=7, geometric_rounds=1, geometric_iters=2)
    cands = [(1, 1, 1), (1, 1, 2), (2, 1, 1), (2, 1, 2), (3, 1, 1)]
    sch, rows = autotune.tune_random(s, [0, 1], plain, gate=0.99, candidates=cands)
    # bad pixels: (1,1,1) 35, (1,1,2) 30, (2,1,1) 15, (2,1,2) 10, (3,1,1) 0 -> only (3,1,1) passes
    assert [round(r[1], 4) for r in rows] == [0.65, 0.70, 0.85, 0.90, 1.0]
    assert (sch.photometric_iters, sch.geometric_rounds, sch.geometric_iters) == (3, 1, 1)
    sch, _ = autotune.tune_random(s, [0], plain, gate=0.85, candidates=cands)
    assert (sch.photometric_iters, sch.geometric_iters) == (2, 1)  # 35 ms, cheaper than (2,1,2) and (3,1,1)
