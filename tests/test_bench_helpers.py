"""bench.py's bookkeeping on the CPU: the algorithmic work per half-step launch that `roofline` and
`roofline_valu` divide by the measured launch time, the accuracy gate's agreement count, and the
rule that confirms a calibration-tuned schedule on every view of the job."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


class _Pm:
    window_radius, window_step = 5, 2


class _Inp:
    """Two views at one level: 20 x 10 and 8 x 6 pixels."""

    pm = _Pm()

    def level_shape(self, v, lv):
        return [(10, 20), (6, 8)][v]


def test_half_step_work_counts_bytes_and_flop_per_launch():
    src = [[1], [0, 0, 0]]  # 1 and 3 sources
    b, f = bench.half_step_work(_Inp(), src, [(0, False), (0, True)])
    # per pixel and half step: 58 + 2 (1 + nsrc) B, + 8 nsrc geometric; a launch covers npix / 2
    photometric = 200 / 2 * (58 + 4) + 48 / 2 * (58 + 8)
    geometric = 200 / 2 * (58 + 4 + 8) + 48 / 2 * (58 + 8 + 24)
    assert abs(b - (photometric + geometric) / 2) < 1e-6
    # 28 FLOP per sample, 15 nsrc WN^2 samples, WN = 6 at radius 5 step 2
    want = (28 * 15 * 1 * 36 * 200 / 2 + 28 * 15 * 3 * 36 * 48 / 2)
    assert abs(f - want) < 1e-6


def test_agreement_counts_depth_within_one_percent_on_the_plain_mask():
    class O:
        def __init__(self, d):
            self.depth = d

    plain = np.array([[10.0, 10.0, 0.0, 20.0]], np.float32)
    raw_plain = np.array([[10.0, 10.0, 5.0, 20.0]], np.float32)
    raw = np.array([[10.09, 10.2, 5.0, 19.9]], np.float32)  # 0.9 %, 2 %, masked, 0.5 %
    n = np.zeros((1, 4, 3), np.float32)
    n[..., 2] = -1
    acc, nrm, tot = bench.agreement([O(plain)], [raw_plain], [raw], [n], [n])
    assert tot == 3 and abs(acc - 2 / 3) < 1e-12 and nrm == 1.0


def test_confirm_tries_calibration_passes_fastest_first_on_every_view():
    rows = [("a", 0.995, 30.0), ("b", 0.98, 10.0), ("c", 0.999, 20.0), ("d", 0.992, 25.0)]
    full = {"a": 0.993, "c": 0.989, "d": 0.991}
    calls = []

    def run(c):
        calls.append(c)
        return full[c], c.upper()

    cand, acc, out, tried = bench.confirm(rows, 0.99, run, "t")
    # b fails calibration; c (20 ms) fails on every view; d (25 ms) passes
    assert calls == ["c", "d"] and cand == "d" and acc == 0.991 and out == "D"
    assert tried == [("c", 0.989), ("d", 0.991)]
    # none confirmed: the most accurate one tried
    full = {"a": 0.98, "c": 0.985, "d": 0.97}
    cand, acc, _, tried = bench.confirm(rows, 0.99, run, "t")
    assert cand == "c" and len(tried) == 3
    # none passing calibration: the most accurate calibration candidate, run once
    calls.clear()
    cand, acc, _, tried = bench.confirm([("b", 0.98, 10.0), ("e", 0.985, 5.0)], 0.99, lambda c: (0.9, None), "t")
    assert cand == "e" and tried == [("e", 0.9)]
