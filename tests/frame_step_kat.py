"""Known-answer scenarios for the tracker frame step (pyrlk_cpu::update,
tracking2/pyrlk_tracker.hpp:831-875, driven by tracker::run, tracker.hpp:73-91),
usable on the oracle tracker and on the GPU tracker alike.

Scenario 1, integer translation with detector_frequency 3 and
filtering_frequency 2: every frame is a crop of one texture shifted by a whole
(2, -2) px per frame, so scale 0 moves by exactly (2, -2) and the 2x2 box level
(scale 1) by exactly (1, -1).  What follows in closed form from the frame step:
  * detection (and its compaction) runs exactly on frames t = 0 mod 3:
    compact_has_run(s) == (t % 3 == 0), and on the other frames the dense
    vector keeps its length (no additions, no compaction: the dead stay);
  * a live particle was born on a detection frame: (t - age + 1) % 3 == 0;
  * every accepted particle of the interior moved by the scale's shift (to
    the tight solver's tolerance), so its speed is the shift and its position
    is its integer birth pixel + (age - 1) * shift (a rare mis-track, a wrong
    local minimum that passes the acceptance tests, is counted and bounded);
  * merging runs only on even frames: on odd frames every accepted particle
    is still alive; on even frames without detection a particle accepted but
    dead was removed by a younger c8 neighbour (merge_trajectories.h:16-27);
  * the sparse image indexes each live particle's pixel with the last index
    at that pixel (kernel_particle_container::move, particle_container.hpp:422-466).

Scenario 2, crafted merges on one scale: three particles with the true speed
are placed around a strong corner before a frame; after it, each is matched
(age + 1) and merge_trajectories removes, for each live particle in index
order, the first c8 neighbour (CPU c8_h order, neighb2d_data.h:80-85) that is
more than 2 frames older with a speed less than 2 px apart.
"""
import numpy as np

SHIFT = (2, -2)  # content motion per frame at scale 0
M = 48           # texture margin (>= frames * |shift|)
NR, NC = 240, 320


def texture(seed, nr, nc):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    t = gaussian_filter(rng.normal(0, 1, (nr, nc)), 2.0)
    t = (t - t.min()) / (t.max() - t.min()) * 255.0
    return np.round(t).astype(np.uint8)


def shifted_frames(nframes, seed=21):
    canvas = texture(seed, NR + 2 * M, NC + 2 * M)
    out = []
    for t in range(nframes):
        r0, c0 = M - t * SHIFT[0], M - t * SHIFT[1]
        out.append(np.ascontiguousarray(canvas[r0:r0 + NR, c0:c0 + NC]))
    return out


def decisions(tr, s):
    return tr.decisions(s) if hasattr(tr, "decisions") else tr.last_decisions(s)


def _interior(p, nr, nc, margin):
    r, c = p["pos_r"], p["pos_c"]
    return (r >= margin) & (r < nr - margin) & (c >= margin) & (c < nc - margin)


# Settings of scenario 1: a tight solver (|delta| < 1e-3, up to 20 steps) so
# that an exact shift is recovered to ~1e-4 px.
CADENCE = dict(detector_frequency=3, filtering_frequency=2, lk_eps=1e-3, lk_max_iters=20)


def check_translation_cadence(tr, nscales=2, nframes=10):
    """Scenario 1 on a tracker made with nscales, profile B, FAST th 20 and
    CADENCE.  Returns counters of what was checked."""
    frames = shifted_frames(nframes)
    seen = dict(accepted=0, interior=0, merged=0, mistracked=0)
    prev_len = [0] * nscales
    for t in range(nframes):
        before = [tr.particles(s) for s in range(nscales)]
        tr.run(frames[t])
        for s in range(nscales):
            nr, nc = NR >> s, NC >> s
            sh = np.array(SHIFT, np.float64) / (1 << s)
            p = tr.particles(s)
            assert tr.compact_has_run(s) == (t % 3 == 0), (t, s)
            if t % 3:
                assert len(p) == len(before[s]) == prev_len[s], (t, s)
            prev_len[s] = len(p)
            live = p[p["age"] > 0]
            age = live["age"].astype(np.int64)
            assert ((t - age + 1) % 3 == 0).all() and (age <= t + 1).all(), (t, s)
            if t == 0:
                continue
            acc, trn, _ = decisions(tr, s)
            assert len(acc) == len(before[s])
            seen["accepted"] += int(acc.sum())
            if t % 3:
                # index-aligned: the matcher's decision and the merge are the only changes
                ok = acc == 1
                assert (p["age"][ok] == before[s]["age"][ok] + 1).all() or t % 2 == 0, (t, s)
                if t % 2 == 1:
                    assert (p["age"][ok] > 0).all(), (t, s)  # no merge on odd frames
                else:
                    dead = np.nonzero(ok & (p["age"] == 0))[0]
                    seen["merged"] += len(dead)
                    # ages after the match, before the merge
                    matched = np.where(ok, before[s]["age"].astype(np.int64) + 1, 0)
                    for i in dead:
                        assert _has_younger_neighbour(p, matched, i), (t, s, i)
                assert (p["age"][~ok] == 0).all(), (t, s)
            # interior particles (at every coarser scale too) tracked the exact
            # shift; a rare mis-track (a wrong local minimum of the coarse solve
            # that the tests still accept) is counted, not failed
            mv = live[(live["age"] >= 2) & _interior(live, nr, nc, 12 << (nscales - 1 - s))]
            off = (np.abs(mv["speed_r"] - sh[0]) >= 0.01) | (np.abs(mv["speed_c"] - sh[1]) >= 0.01)
            seen["interior"] += len(mv)
            seen["mistracked"] += int(off.sum())
            mv = mv[~off]
            k = mv["age"].astype(np.float64) - 1.0
            br = mv["pos_r"].astype(np.float64) - k * sh[0]
            bc = mv["pos_c"].astype(np.float64) - k * sh[1]
            # a particle that mis-tracked once keeps its offset: birth pixels of
            # the rest are integers within the accumulated solver tolerance
            far = (np.abs(br - np.round(br)) >= 0.005 * np.maximum(k, 1)) | \
                (np.abs(bc - np.round(bc)) >= 0.005 * np.maximum(k, 1))
            seen["mistracked"] += int(far.sum())
            # the sparse image: the last index of each live particle's pixel
            sp = tr.sparse(s)
            idx = np.nonzero(p["age"] > 0)[0]
            pr, pc = p["pos_r"][idx].astype(np.int64), p["pos_c"][idx].astype(np.int64)
            got = sp[pr, pc]
            assert (got >= idx).all(), (t, s)
            assert np.array_equal(p["pos_r"][got].astype(np.int64), pr) and \
                np.array_equal(p["pos_c"][got].astype(np.int64), pc), (t, s)
    return seen


C8 = [(-1, 1), (0, 1), (1, 1), (-1, 0), (1, 0), (-1, -1), (0, -1), (1, -1)]


def _has_younger_neighbour(p, age, i):
    """A particle matched this frame (age > 0 before the merge) sits at a c8
    neighbour of particle i and is more than 2 frames younger."""
    r, c = int(p["pos_r"][i]), int(p["pos_c"][i])
    tr_, tc = p["pos_r"].astype(np.int64), p["pos_c"].astype(np.int64)
    for dr, dc in C8:
        nb = np.nonzero((tr_ == r + dr) & (tc == c + dc) & (age > 0))[0]
        if any(age[i] > age[j] + 2 for j in nb):
            return True
    return False


def strongest_corner(orc, frame, margin=24):
    sal = orc.fast_saliency(frame, None, 9, 20)
    inner = np.zeros_like(sal)
    inner[margin:-margin, margin:-margin] = sal[margin:-margin, margin:-margin]
    r, c = np.unravel_index(int(np.argmax(inner)), inner.shape)
    return int(r), int(c)


def crafted(dtype, corner, layout):
    """Particles (name, offset from the corner, age) at pixel centres with the
    true speed."""
    out = np.zeros(len(layout), dtype)
    for k, (_, (dr, dc), age) in enumerate(layout):
        out[k]["pos_r"] = corner[0] + dr + 0.5
        out[k]["pos_c"] = corner[1] + dc + 0.5
        out[k]["speed_r"], out[k]["speed_c"] = SHIFT
        out[k]["age"] = age
    return out


# (layout, expected ages after one frame): A old, B young beside it, C older
# still at B's first c8 neighbour.
MERGE_CASES = [
    # B removes A (c8 offset (0,-1) from B)
    ([("A", (0, 0), 10), ("B", (0, 1), 3)], [0, 4]),
    # the same in the other index order
    ([("B", (0, 1), 3), ("A", (0, 0), 10)], [4, 0]),
    # B meets C first in c8_h order ((-1,1) comes before (0,-1)) and stops there
    ([("A", (0, 0), 10), ("B", (0, 1), 3), ("C", (-1, 2), 12)], [11, 4, 0]),
    # A only 2 frames older than B after the match: nothing is removed
    ([("A", (0, 0), 5), ("B", (0, 1), 3)], [6, 4]),
    # a removed particle does not act: C (removed by B first) would have removed A
    ([("B", (0, 1), 1), ("C", (-1, 2), 12), ("A", (0, 0), 20)], [2, 0, 21]),
]


def check_crafted_merges(tr_factory, orc, particle_dtype):
    """Scenario 2: for each case a fresh single-scale tracker (detector only on
    frame 0, merge every frame) runs frame 0, gets the crafted particles, runs
    frame 1."""
    frames = shifted_frames(2, seed=33)
    corner = strongest_corner(orc, frames[0])
    for layout, want in MERGE_CASES:
        tr = tr_factory()
        tr.run(frames[0])
        tr.set_particles(0, crafted(particle_dtype, corner, layout))
        tr.run(frames[1])
        acc, trn, _ = decisions(tr, 0)
        assert acc.tolist() == [1] * len(layout), (layout, acc)
        assert np.abs(trn - np.array(SHIFT, np.float32)).max() < 0.01, (layout, trn)
        p = tr.particles(0)
        assert p["age"].tolist() == want, (layout, p["age"].tolist(), want)
