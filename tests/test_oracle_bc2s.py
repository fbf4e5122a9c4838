"""Pins of the oracle's bc2s path (the reference sample's default tracker):

  * 1-2-1 blur: impulse and step-edge responses worked out by hand, constants
    kept, and a numpy restatement on edge-padded arrays;
  * integer Gaussian taps from the closed form (FP64) and its impulse
    response;
  * contrast: the responses to an impulse and to a step edge (by hand);
  * descriptor: the reference's radius-3 ring table (neighb2d_data.h:184-189)
    as data, every other point at scale 1 and twice the offset at scale 2;
  * distance: zero at the descriptor's own pixel, constant images by hand,
    scale 2 blind to s1;
  * descent: a blob recovered from offset predictions (sign of the
    displacement), the scale-1 start distance of the scale-2 descent with its
    c8_h first neighbour, and an independent Python restatement of the scan
    order on random landscapes;
  * frame step: an integer-shift video (speeds equal the shift), a static
    video (positions and speeds frozen, ages counting), the flow statistics
    and multiscale counts recomputed from the particles, and the filter's
    post-condition.
"""
import math

import numpy as np
import pytest

import oracle

# circle_r3_h (cuimg/neighb2d_data.h:184-189), (row, col)
CIRCLE_R3 = [(-3, 0), (-3, 1), (-2, 2), (-1, 3), (0, 3), (1, 3), (2, 2), (3, 1),
             (3, 0), (3, -1), (2, -2), (1, -3), (0, -3), (-1, -3), (-2, -2), (-3, -1)]
# c8_h and c8_it_h (cuimg/neighb2d_data.h:80-85, 162-180)
C8 = [(-1, 1), (0, 1), (1, 1), (-1, 0), (1, 0), (-1, -1), (0, -1), (1, -1)]
C8_IT = [(6, 3), (0, 3), (0, 5), (2, 5), (2, 7), (4, 7), (4, 1), (6, 1), (0, 0)]


def test_blur121_impulse_and_edges(orc):
    img = np.zeros((7, 7), np.uint8)
    img[3, 3] = 255
    out = orc.blur121(img)
    want = np.zeros((7, 7), np.uint8)
    # rows: 510/4 = 127 at the centre, 255/4 = 63 beside it; then columns
    want[3, 3] = 63            # 2*127/4
    want[2, 3] = want[4, 3] = 31   # 127/4
    want[3, 2] = want[3, 4] = 31   # 2*63/4
    want[2, 2] = want[2, 4] = want[4, 2] = want[4, 4] = 15  # 63/4
    assert np.array_equal(out, want)
    # a step edge along the rows, with the clamped ends
    step = np.tile(np.array([0, 0, 4, 4], np.uint8), (5, 1))
    assert np.array_equal(orc.blur121(step), np.tile(np.array([0, 1, 3, 4], np.uint8), (5, 1)))
    assert np.array_equal(orc.blur121(step.T.copy()), np.tile(np.array([0, 1, 3, 4], np.uint8), (5, 1)).T)
    for v in (0, 1, 77, 255):
        assert np.array_equal(orc.blur121(np.full((6, 9), v, np.uint8)), np.full((6, 9), v, np.uint8))


def blur121_numpy(img):
    a = np.pad(img.astype(np.int32), 1, mode="edge")
    t = (2 * a[:, 1:-1] + a[:, :-2] + a[:, 2:]) // 4
    return ((2 * t[1:-1] + t[:-2] + t[2:]) // 4).astype(np.uint8)


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (6, 1), (31, 17), (64, 80)])
def test_blur121_numpy_restatement(orc, shape):
    img = np.random.default_rng(shape[0] * 100 + shape[1]).integers(0, 256, shape).astype(np.uint8)
    assert np.array_equal(orc.blur121(img), blur121_numpy(img))


def test_integer_gaussian_taps_and_impulse(orc):
    for sigma, hw in ((2.0, 3), (1.0, 1), (1.5, 4), (3.0, 9)):
        want = [int(100 * math.exp(-i * i / (2 * sigma * sigma)) / (sigma * math.sqrt(2 * math.pi)))
                for i in range(-hw, hw + 1)]
        assert list(orc.gaussian_int_taps(sigma, hw)) == want
    assert list(orc.gaussian_int_taps(2.0, 3)) == [6, 12, 17, 19, 17, 12, 6]
    assert list(orc.gaussian_int_taps(1.0, 1)) == [24, 39, 24]
    img = np.zeros((9, 9), np.uint8)
    img[4, 4] = 255
    out = orc.gaussian_int(img, 1.0, 1)
    col = [255 * 24 // 87, 255 * 39 // 87, 255 * 24 // 87]  # 70, 114, 70
    want = np.zeros((9, 9), np.uint8)
    for i, cv in enumerate(col):
        for j, w in enumerate((24, 39, 24)):
            want[3 + i, 3 + j] = cv * w // 87
    assert np.array_equal(out, want)
    assert np.array_equal(orc.gaussian_int(np.full((5, 7), 200, np.uint8), 2.0, 3), np.full((5, 7), 200, np.uint8))


def test_contrast_known_answers(orc):
    img = np.zeros((12, 12), np.uint8)
    img[:, 5:] = 10
    con = orc.contrast(img)
    want = np.zeros((12, 12), np.uint32)
    want[:, 3:7] = 30  # three of the four pairs straddle the edge
    assert np.array_equal(con, want)
    imp = np.zeros((13, 13), np.uint8)
    imp[6, 6] = 100
    con = orc.contrast(imp)
    want = np.zeros((13, 13), np.uint32)
    for dr, dc in ((0, 2), (0, -2), (2, 0), (-2, 0), (2, 2), (-2, -2), (2, -2), (-2, 2)):
        want[6 + dr, 6 + dc] = 100
    assert np.array_equal(con, want)


def test_descriptor_ring_offsets(orc):
    r_, c_ = np.mgrid[0:40, 0:40]
    s1 = ((r_ * 13 + c_ * 7) % 251).astype(np.uint8)
    s2 = ((r_ * 5 + c_ * 11 + 3) % 251).astype(np.uint8)
    for r, c in ((20, 20), (8, 31), (6, 6)):
        f = orc.bc2s_feature(s1, s2, r, c)
        for i in range(8):
            dr, dc = CIRCLE_R3[2 * i]
            assert f[i] == s1[r + dr, c + dc]
            assert f[8 + i] == s2[r + 2 * dr, c + 2 * dc]
    # next to the edge the samples clamp (the images' borders are clamp-filled)
    f = orc.bc2s_feature(s1, s2, 0, 0)
    assert f[0] == s1[0, 0] and f[8] == s2[0, 0]


def test_distance_known_answers(orc):
    s1 = np.full((30, 30), 10, np.uint8)
    s2 = np.full((30, 30), 20, np.uint8)
    f = np.zeros(16, np.uint8)
    assert orc.bc2s_distance(s1, s2, f, 15, 15, 1) == 8 * 10 + 8 * 20
    assert orc.bc2s_distance(s1, s2, f, 15, 15, 2) == 8 * 20
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (40, 40)).astype(np.uint8)
    b = rng.integers(0, 256, (40, 40)).astype(np.uint8)
    f = orc.bc2s_feature(a, b, 20, 17)
    assert orc.bc2s_distance(a, b, f, 20, 17, 1) == 0
    assert orc.bc2s_distance(a, b, f, 20, 17, 2) == 0
    d2 = orc.bc2s_distance(a, b, f, 21, 18, 2)
    assert orc.bc2s_distance(255 - a, b, f, 21, 18, 2) == d2  # scale 2 ignores s1


def blob(n, r0, c0, sigma):
    r, c = np.mgrid[0:n, 0:n]
    return (250 * np.exp(-((r - r0) ** 2 + (c - c0) ** 2) / (2 * sigma * sigma))).astype(np.uint8)


def test_descent_recovers_a_blob(orc):
    img = blob(64, 30, 33, 5.0)
    s1, s2 = orc.bc2s_images(img)
    f = orc.bc2s_feature(s1, s2, 30, 33)
    assert orc.bc2s_match(s1, s2, f, (30, 33)) == ((0, 0), 0.0)
    for dr, dc in ((1, 1), (-2, 1), (0, -3), (2, 2)):
        d, dist = orc.bc2s_match(s1, s2, f, (30 + dr, 33 + dc))
        assert d == (-dr, -dc) and dist == 0.0, (dr, dc, d, dist)


def test_scale2_descent_starts_from_the_scale1_distance(orc):
    """With s2 flat, every neighbour's scale-2 distance beats the starting
    scale-1 distance (which adds the s1 ring): the descent moves once, to the
    first c8_h neighbour (-1, +1), and stops there."""
    rng = np.random.default_rng(3)
    s1 = rng.integers(0, 256, (40, 40)).astype(np.uint8)
    s2 = np.full((40, 40), 90, np.uint8)
    f = np.concatenate([np.zeros(8, np.uint8), np.full(8, 80, np.uint8)])
    d, dist = orc.bc2s_match(s1, s2, f, (20, 20), scale=2)
    assert d == C8[0] and dist == 80.0


def descent_py(s1, s2, f, pred, scale):
    """gradient_descent_match restated from the reference's loop (tables above)."""
    nr, nc = s1.shape
    s1p, s2p = np.pad(s1.astype(int), 3, mode="edge"), np.pad(s2.astype(int), 6, mode="edge")
    f = f.astype(int)

    def dist(r, c, sc):
        d = sum(abs(s2p[r + 6 + 2 * CIRCLE_R3[2 * i][0], c + 6 + 2 * CIRCLE_R3[2 * i][1]] - f[8 + i]) for i in range(8))
        if sc == 1:
            d += sum(abs(s1p[r + 3 + CIRCLE_R3[2 * i][0], c + 3 + CIRCLE_R3[2 * i][1]] - f[i]) for i in range(8))
        return d

    def inside(r, c):
        return 0 <= r < nr and 0 <= c < nc

    p = tuple(pred)
    if not inside(*p):
        return p, 999999.0
    best, bd, bi = p, dist(*p, 1), 8
    for _ in range(10):
        start, end = C8_IT[bi]
        order = [start]
        k = (start + 1) % 8
        while k != end:
            order.append(k)
            k = (k + 1) % 8
        for k in order:
            n = (p[0] + C8[k][0], p[1] + C8[k][1])
            if inside(*n):
                dn = dist(*n, scale)
                if dn < bd:
                    best, bd, bi = n, dn, k
        if best == p:
            break
        p = best
    return (best[0] - pred[0], best[1] - pred[1]), float(bd)


@pytest.mark.parametrize("seed", range(4))
def test_descent_python_restatement(orc, seed):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, (48, 56)).astype(np.uint8)
    img = np.clip(blur121_numpy(blur121_numpy(img)).astype(int) * 2 - 128, 0, 255).astype(np.uint8)
    s1, s2 = orc.bc2s_images(img)
    for _ in range(60):
        q = rng.integers(0, 48), rng.integers(0, 56)
        f = orc.bc2s_feature(s1, s2, *q)
        pred = (int(q[0] + rng.integers(-4, 5)), int(q[1] + rng.integers(-4, 5)))
        for scale in (1, 2):
            assert orc.bc2s_match(s1, s2, f, pred, scale) == descent_py(s1, s2, f, pred, scale)
        m1, _ = descent_py(s1, s2, f, pred, 2)
        m2, d2 = descent_py(s1, s2, f, (pred[0] + m1[0], pred[1] + m1[1]), 1)
        assert orc.bc2s_match(s1, s2, f, pred, 0) == ((m1[0] + m2[0], m1[1] + m2[1]), d2)


def shift_video(seed, nr, nc, nframes, dr, dc):
    rng = np.random.default_rng(seed)
    pad = 4 + max(abs(dr), abs(dc)) * nframes
    big = rng.integers(0, 256, (nr + 2 * pad, nc + 2 * pad)).astype(np.uint8)
    big = blur121_numpy(blur121_numpy(big))
    big = np.clip((big.astype(int) - 128) * 3 + 128, 0, 255).astype(np.uint8)
    return np.stack([big[pad - f * dr:pad - f * dr + nr, pad - f * dc:pad - f * dc + nc] for f in range(nframes)])


def test_frame_step_integer_shift(orc):
    video = shift_video(2, 120, 160, 8, 2, -1)
    t = orc.tracker(120, 160, oracle.profile_params("C", 2, 20.0, detector_frequency=4))
    for f in range(8):
        t.run(video[f])
    p = t.particles(0)
    live = p[p["age"] > 1]
    assert len(live) > 200
    assert ((live["speed_r"] == 2) & (live["speed_c"] == -1)).mean() > 0.9
    # positions and speeds are integers
    for k in ("pos_r", "pos_c", "speed_r", "speed_c"):
        assert np.array_equal(p[k], np.round(p[k]))


def filter_kills(q, flow, msc):
    """filter_bad_particles_kernel's removal rule for one particle."""
    R, C = int(q["pos_r"]) // 8, int(q["pos_c"]) // 8
    m = msc[R, C]
    if q["age"] > 2 and m > 0:
        if math.hypot(q["speed_r"] - flow[R, C, 0], q["speed_c"] - flow[R, C, 1]) > 10.0:
            return True
    alone = all(not (0 <= R + dr < msc.shape[0] and 0 <= C + dc < msc.shape[1]) or msc[R + dr, C + dc] == 0
                for dr, dc in C8)
    return alone and m == 1


def test_frame_step_static_video(orc):
    """Nothing moves; survivors age by one per frame; a particle only dies by
    the filter (no detector here, and merges need different ages): in a
    static scene that is a cell without flow (NO_FLOW is 500 px away from the
    zero speed) once the particle is older than 2, or an isolated one."""
    frame = shift_video(4, 100, 120, 1, 0, 0)[0]
    t = orc.tracker(100, 120, oracle.profile_params("C", 2, 15.0, detector_frequency=100))
    t.run(frame)
    before = t.particles(0).copy()
    assert (before["age"] > 0).sum() > 50
    died = 0
    for k in range(4):
        t.run(frame)
        p = t.particles(0)
        flow, msc = t.flow(0)
        assert len(p) == len(before)
        for f in ("pos_r", "pos_c"):
            assert np.array_equal(p[f], before[f])
        assert not p["speed_r"].any() and not p["speed_c"].any()
        was = before["age"] > 0
        alive = p["age"] > 0
        assert np.array_equal(p["age"][alive], before["age"][alive] + 1)
        for i in np.nonzero(was & ~alive)[0]:
            q = p[i].copy()
            q["age"] = before["age"][i] + 1
            assert filter_kills(q, flow, msc), i
            died += 1
        before = p.copy()
    assert died > 0 and (before["age"] > 0).sum() > 20


def test_flow_statistics_and_filter_invariants(orc):
    video = shift_video(7, 96, 128, 7, 1, 1)
    t = orc.tracker(96, 128, oracle.profile_params("C", 2, 12.0, detector_frequency=3))
    for f in range(7):
        t.run(video[f])
        stats = []
        for s in range(2):
            flow, msc = t.flow(s)
            sp = t.sparse(s)
            dec = t.last_decisions(s)[0]
            nr, nc = sp.shape
            cr, cc = flow.shape[:2]
            assert (cr, cc) == (-(-nr // 8), -(-nc // 8))
            stats.append((flow, msc, cr, cc))
        # counts: cells of the finest scale from its own count plus the coarser one
        flow0, msc0, cr, cc = stats[0]
        flow1, msc1, cr1, cc1 = stats[1]
        for R in range(cr):
            for C in range(cc):
                up = msc1[R // 2, C // 2] if (R // 2 < cr1 and C // 2 < cc1) else 0
                assert msc0[R, C] >= up
        # the filter's post-condition on the last state
        for s in range(2):
            flow, msc, cr_, cc_ = stats[s]
            p = t.particles(s)
            for q in p[p["age"] > 2]:
                R, C = int(q["pos_r"]) // 8, int(q["pos_c"]) // 8
                if msc[R, C] > 0:
                    d = math.hypot(q["speed_r"] - flow[R, C, 0], q["speed_c"] - flow[R, C, 1])
                    assert d <= 10.0 + 1e-6
