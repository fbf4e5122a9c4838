"""Known answers for the oracle's frame step (tests/frame_step_kat.py): the
detector / merge cadence, ages, positions and the sparse image on an exact
integer-translation video, and crafted merges whose outcome follows from the
reference's merge loop by hand."""
import oracle

import frame_step_kat as kat


def test_oracle_translation_cadence_known_answer(orc):
    p = oracle.profile_params("B", 2, 20.0, **kat.CADENCE)
    seen = kat.check_translation_cadence(orc.tracker(kat.NR, kat.NC, p))
    print(seen)
    assert seen["interior"] > 20000 and seen["accepted"] > 20000 and seen["merged"] > 0
    assert seen["mistracked"] <= 0.002 * seen["interior"]


def test_oracle_crafted_merges_known_answer(orc):
    p = oracle.profile_params("B", 1, 20.0, detector_frequency=1000, filtering_frequency=1)
    kat.check_crafted_merges(lambda: orc.tracker(kat.NR, kat.NC, p), orc, oracle.PARTICLE_DTYPE)
