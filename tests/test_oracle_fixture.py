"""The oracle's frame step against the committed end-to-end fixture
(tests/golden/frame_step_320x240.json, written by
tests/golden/make_frame_step_fixture.py): a regression pin of every frame's
dense vector, sparse image, saliency and matches, for both matcher profiles
and both arithmetic readings and for the bc2s strategy (with its descriptors,
flow and multiscale counts), and of the synthetic video itself."""
import json
import os

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "frame_step_320x240.json")


@pytest.fixture(scope="module")
def fixture():
    with open(GOLDEN) as f:
        return json.load(f)


def test_video_is_the_fixture_video(fixture):
    import golden.make_frame_step_fixture as m
    assert m.digest(m.video()) == fixture["video"]["digest"]


def test_oracle_bc2s_frame_steps_match_fixture(orc, fixture):
    import golden.make_frame_step_fixture as m
    t = orc.tracker(m.NR, m.NC, oracle.profile_params("C", m.NSCALES, m.TH, detector_frequency=m.C_DET_FREQ))
    v = m.video()
    for f, want in enumerate(fixture["runs"]["C/exact"]):
        t.run(v[f])
        assert m.state_bc2s(t, m.NSCALES) == want, f


@pytest.mark.parametrize("run", ["B/raster", "B/gpu", "R/raster", "R/gpu"])
def test_oracle_frame_steps_match_fixture(orc, fixture, run):
    import golden.make_frame_step_fixture as m
    profile, arith = run.split("/")
    ws = 7 if profile == "B" else 11
    lanes = oracle.gpu_sum_lanes(ws) if arith == "gpu" else 0
    t = orc.tracker(m.NR, m.NC, oracle.profile_params(profile, m.NSCALES, m.TH, sum_lanes=lanes))
    v = m.video()
    for f, want in enumerate(fixture["runs"][run]):
        t.run(v[f])
        assert m.state(t, m.NSCALES) == want, (run, f)
