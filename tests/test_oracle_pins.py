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
