"""Plants plausible mistakes in the CPU oracle (on the CPU) and checks that the pins of
tests/test_oracle_kats.py catch each of them: the pins would not pass a wrong restatement.

Each mutant is the oracle compiled with -DPMO_MUT=n (see the hooks in oracle/pm_oracle.c):
a wrong RNG tag shift, a wrong emission exponent, a transposed rotation, swapped bilinear axes,
dropped terms in the variance and covariance, a truncated exp polynomial, swapped homography
operands, a sign slip in the triangulation prior and a wrong FNV prime.
"""
import os
import subprocess
import sys

import pytest

import oracle

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import test_oracle_kats as kats  # noqa: E402

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
