"""The oracle's pins have teeth: each plausible mistake, planted in a
separate build of the oracle (PLANT_* macros in oracle/flct_oracle.c),
makes at least one pinning check fail.  CPU only."""
import os
import subprocess

import pytest

import oracle_lib as O
import test_oracle as T

SRC = os.path.join(O.ROOT, "oracle", "flct_oracle.c")

PLANTS = {
    "PLANT_FILTER_ONCE": [T.test_flct_reproduces_reference_fixture],
    "PLANT_TIE_LAST": [T.test_peak_edge_and_zero_rules],
    "PLANT_FIT_SIGN": [T.test_peak_fit_on_sampled_quadratic,
                       T.test_flct_reproduces_reference_fixture],
    "PLANT_NO_MEAN": [T.test_flct_reproduces_reference_fixture],
    "PLANT_NO_SWAP": [T.test_flct_reproduces_reference_fixture,
                      T.test_uniform_shift_recovers_known_velocity],
    "PLANT_NYQUIST_NEG": [T.test_make_freq_closed_form],
    "PLANT_HIRES_SCALE": [T.test_hires_peak_on_sampled_quadratic],
    # -b: offset d instead of d/2, windows moved the wrong way, means
    # recomputed over the moved windows, only image 2's window moved
    "PLANT_BIAS_FULL_D": [T.test_bias_matches_direct_sum_definition,
                          T.test_bias_exact_on_a_moved_integer_blob],
    "PLANT_BIAS_SIGN": [T.test_bias_matches_direct_sum_definition,
                        T.test_bias_exact_on_a_moved_integer_blob],
    "PLANT_BIAS_REMEAN": [T.test_bias_matches_direct_sum_definition,
                          T.test_bias_zero_offset_is_the_plain_pass],
    "PLANT_BIAS_ONE_IMAGE": [T.test_bias_matches_direct_sum_definition,
                             T.test_bias_exact_on_a_moved_integer_blob,
                             T.test_bias_swap_negates_velocity],
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_is_caught(plant, tmp_path, golden):
    so = str(tmp_path / f"lib_{plant}.so")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", f"-D{plant}",
                    "-o", so, SRC, "-lm"], check=True)
    with O.using(so):
        for check in PLANTS[plant]:
            args = [golden] if "golden" in check.__code__.co_varnames[:check.__code__.co_argcount] else []
            with pytest.raises(AssertionError):
                check(*args)
