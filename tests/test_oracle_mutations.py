"""Plants plausible mistakes in the CPU oracle (on the CPU) and checks that the pins of
tests/test_oracle_kats.py catch each of them: the pins would not pass a wrong restatement.

Each mutant is the oracle compiled with -DPMO_MUT=n (see the hooks in oracle/pm_oracle.c):
a wrong RNG tag shift, a wrong emission exponent, a transposed rotation, swapped bilinear axes,
dropped terms in the variance and covariance, a truncated exp polynomial, swapped homography
operands, a sign slip in the triangulation prior and a wrong FNV prime; and, in the decision rules
pinned by tests/test_oracle_rules.py: a good-count edge (>= n1), a bad-count edge (<= n2), a
dropped doubling of the previous view, a dropped prev_view_weight, later-index ties for the
hypothesis, the view, the propagation region and the refinement, fallback weights left negative,
one perturbation try instead of four, a top-K that forgets to shift its third value, a top-K that
divides by K instead of min(K, n), a triangulation equality that fails, a zero incidence that
passes and a count threshold without min(., nsrc).
"""
import os
import subprocess
import sys

import pytest

import oracle

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import test_oracle_kats as kats  # noqa: E402
import test_oracle_rules as rules  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# mutation id -> the pin that must reject it
CASES = {
    1: kats.test_rng_state_definition,
    2: kats.test_math_reference_values,
    3: kats.test_level_pair_homography_projects_plane_points,
    4: kats.test_ncc_fp64_brute_force,
    5: kats.test_ncc_fp32_agrees_with_fp64,
    6: kats.test_ncc_closed_forms,
    7: kats.test_expf_matches_libm,
    8: kats.test_ncc_fp32_agrees_with_fp64,
    9: kats.test_math_reference_values,
    10: kats.test_fnv1a_published_vectors,
    11: rules.test_vote_edges,
    12: rules.test_vote_edges,
    13: rules.test_view_weight_previous_view_and_prior,
    14: rules.test_view_weight_previous_view_and_prior,
    15: rules.test_hypothesis_and_view_ties_go_to_the_earlier_index,
    16: rules.test_hypothesis_and_view_ties_go_to_the_earlier_index,
    17: rules.test_prior_weighted_fallback,
    18: rules.test_normal_perturbation_halves_up_to_four_tries,
    19: rules.test_region_pick_offsets_order_and_ties,
    20: rules.test_refinement_acceptance,
    21: rules.test_topk_mean,
    22: rules.test_topk_mean,
    23: rules.test_readout_source_tests,
    24: rules.test_readout_source_tests,
    25: rules.test_count_threshold,
}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    src = os.path.join(ROOT, "oracle", "pm_oracle.c")
    fsrc = os.path.join(ROOT, "oracle", "fusion_oracle.c")
    psrc = os.path.join(ROOT, "oracle", "post_oracle.c")
    ssrc = os.path.join(ROOT, "oracle", "scene_oracle.c")
    paths = {}
    procs = []
    for m in CASES:
        out = str(d / f"libpm_oracle_mut{m}.so")
        cmd = ["gcc", "-O1", "-march=x86-64-v3", "-ffp-contract=off", "-fno-math-errno", "-fopenmp", "-fPIC", "-shared",
               f"-DPMO_MUT={m}", "-o", out, src, fsrc, psrc, ssrc, "-lm"]
        procs.append(subprocess.Popen(cmd))
        paths[m] = out
    for p in procs:
        assert p.wait() == 0
    return paths


@pytest.mark.parametrize("mut", sorted(CASES))
def test_pin_rejects_planted_mistake(mut, mutants, monkeypatch):
    lib = oracle.load_library(mutants[mut])
    monkeypatch.setattr(oracle, "_lib", lib)
    with pytest.raises(AssertionError):
        CASES[mut](oracle)


def test_unmutated_oracle_passes_the_same_pins():
    for f in set(CASES.values()):
        f(oracle)
