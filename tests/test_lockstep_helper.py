"""CPU checks of the lockstep comparison itself (tests/lockstep.py), run with two
oracle trackers in place of the oracle and the GPU: identical trackers give
exact frames only; a tracker that matched one particle differently gives
partial frames that pass; a planted mistake anywhere else fails."""
import numpy as np
import pytest

import cuimg_amd as cu
import lockstep
import oracle


class AsGpu:
    """An oracle tracker with the GPU tracker's accessor names; `age_bump`
    plants a mistake: after each frame particle `age_bump` of scale 0 reads
    one frame older."""

    def __init__(self, t, age_bump=None):
        self.t = t
        self.age_bump = age_bump

    def __getattr__(self, k):
        return getattr(self.t, k)

    def decisions(self, s):
        return self.t.last_decisions(s)

    def particles(self, s):
        p = self.t.particles(s)
        if self.age_bump is not None and s == 0 and len(p) > self.age_bump:
            p = p.copy()
            p["age"][self.age_bump] += 1
        return p


def _run(orc, nframes, shift_one=False, age_bump=None):
    video = cu.synth_video(0xC0FFEE01, 240, 320, nframes, 0)
    prm = oracle.profile_params("B", 3)
    o, g = orc.tracker(240, 320, prm), orc.tracker(240, 320, prm)
    gg = AsGpu(g, age_bump)
    stats = lockstep.new_stats()
    for f in range(nframes):
        pre = [o.particles(s) for s in range(3)]
        pre_g = [p.copy() for p in pre]
        if shift_one and f > 0:
            # one particle starts the frame one pixel off: its match differs
            k = int(np.argmax(pre_g[0]["age"]))
            pre_g[0]["pos_r"][k] += 1.0
        for s in range(3):
            g.set_particles(s, pre_g[s])
        o.run(video[f])
        g.run(video[f])
        lockstep.compare_frame(o, gg, 3, stats, pre, pre_g=pre_g)
    return stats


def test_identical_trackers_give_exact_frames(orc):
    st = _run(orc, 5)
    assert st["frames_exact"] == 5 and st["frames_partial"] == 0
    assert st["within"] == st["tracked"] > 0


def test_one_disagreeing_particle_gives_passing_partial_frames(orc):
    st = _run(orc, 5, shift_one=True)
    assert st["frames_partial"] >= 3, st
    assert st["partial_checked"] > 100 * st["partial_skipped"] > 0, st


def test_planted_age_mistake_is_caught(orc):
    with pytest.raises(AssertionError):
        _run(orc, 4, age_bump=7)
    with pytest.raises(AssertionError):
        _run(orc, 4, shift_one=True, age_bump=7)
