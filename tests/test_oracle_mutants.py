"""The oracle's pinning checks catch plausible mistakes.

adsr_oracle.c built with -DORC_MUTANT=n plants one mistake; each build is
loaded in place of the oracle and the checks of test_oracle.py that pin
that function must fail on it, while they pass on the oracle itself.
"""
import os
import subprocess

import pytest

import oracle
import test_oracle as T
import test_oracle_quant as TQ
import test_oracle_score as TS

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "oracle", "adsr_oracle.c")
SRC_SCORE = os.path.join(os.path.dirname(HERE), "oracle", "adsr_oracle_score.c")

# planted mistake -> (what, the checks that pin it)
MUTANTS = {
    1: ("convolution tap index off by one",
        [T.test_convolution_constant_row_closed_form,
         T.test_convolution_impulse_is_the_mirrored_kernel]),
    2: ("convolution centre term dropped",
        [T.test_convolution_constant_row_closed_form,
         T.test_convolution_impulse_is_the_mirrored_kernel]),
    3: ("score ln(x) instead of ln(x + 1)", [T.test_score_kat_and_closed_form]),
    4: ("local maxima with a non-strict rise",
        [T.test_local_maxima_brute_force_and_stable_order]),
    5: ("binary search returns the first equal element",
        [T.test_rust_bsearch_last_equal_and_insertion_point]),
    6: ("Gaussian exponent without the 1/2", [T.test_kernel_closed_form_and_symmetry]),
    # scoring (adsr_oracle_score.c); "features" is the float64 feature check
    # (the reference's correlation KATs use y = 2x and y = -x, where this
    # mistake is clamped back to +-1; random vectors against numpy catch it)
    7: ("correlation with one variance twice",
        [TS.test_correlation_median_std_against_numpy, "features"]),
    8: ("ranked mask in ascending order", [TS.test_crate_utils_kats, "features"]),
    9: ("hyperscore without the factorial terms", ["features"]),
    10: ("document frequency not floored at 1", [TS.test_idf_kats]),
    11: ("unweighted running m/z", [TS.test_dense_xic_mz_kats]),
    12: ("profiles normalised by the centre cycle only", ["features"]),
    # quantification (orc_quantify); "quantify" is its float64 check
    13: ("observed intensity without the last cycle", ["quantify"]),
    14: ("median over every profile, the all-zero ones too", ["quantify"]),
    15: ("rt_observed of the first cycle", ["quantify"]),
}


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    out = tmp_path_factory.mktemp("mutants")
    paths = {}
    for m in MUTANTS:
        so = str(out / f"liboracle_m{m}.so")
        subprocess.run(
            ["gcc", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11",
             "-shared", f"-DORC_MUTANT={m}", "-o", so, SRC, SRC_SCORE, "-lm",
             "-lpthread"],
            check=True)
        paths[m] = so
    return paths


def _run_check(check, run0):
    # handles built by the library loaded now (a mutant or the oracle)
    if check == "features":
        TS.test_score_features_against_float64_implementation(
            run0, (oracle.Dia(run0["dia"]), oracle.Lib(run0["lib"])))
    elif check == "quantify":
        TQ.test_quantify_against_float64_implementation(
            run0, (oracle.Dia(run0["dia"]), oracle.Lib(run0["lib"])))
    else:
        check()


@pytest.mark.parametrize("m", sorted(MUTANTS), ids=lambda m: MUTANTS[m][0])
def test_planted_mistake_is_caught(builds, m, run0):
    prev = oracle.use_library(builds[m])
    try:
        for check in MUTANTS[m][1]:
            with pytest.raises(AssertionError):
                _run_check(check, run0)
    finally:
        oracle.use_library(prev)


def test_checks_pass_on_the_oracle(run0):
    for _, checks in MUTANTS.values():
        for check in checks:
            _run_check(check, run0)
