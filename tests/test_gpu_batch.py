"""Batched frame step: every stream of a batch must equal its own tracker.

A batch runs one fixed sequence of launches for all (stream, scale) jobs and
replays it from a captured hipGraph; a tracker on its own is a batch of one.
Both are compared bit-exactly (particles, sparse index image, saliency, matches,
decisions) against trackers stepped with plain launches, frame by frame."""
import numpy as np
import pytest

import cuimg_amd as cu

pytestmark = pytest.mark.gpu

SHAPES = [(120, 160), (144, 256), (200, 280), (96, 128), (240, 320)]


@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


def stream_params(i, profile="B"):
    kw = {}
    if i == 2:
        kw = dict(detector_frequency=2, filtering_frequency=3)
    if i == 3:
        kw = dict(with_merge=0)
    return cu.profile_params(profile, 3 if i % 2 == 0 else 2, [20.0, 12.0, 16.0, 8.0, 25.0][i], **kw)


def state(t, L):
    out = []
    for s in range(L):
        out.append((t.particles(s), t.sparse(s), t.saliency(s), t.compact_has_run(s),
                    t.matches(s) if t.compact_has_run(s) else None, t.decisions(s)))
    return out


def assert_same(a, b, tag):
    for s, (x, y) in enumerate(zip(a, b)):
        px, py = x[0], y[0]
        assert len(px) == len(py), (tag, s)
        for k in px.dtype.names:
            assert np.array_equal(px[k], py[k]), (tag, s, k)
        assert np.array_equal(x[1], y[1]), (tag, s, "sparse")
        assert np.array_equal(x[2], y[2]), (tag, s, "saliency")
        assert x[3] == y[3], (tag, s, "compact_has_run")
        if x[4] is not None:
            assert np.array_equal(x[4], y[4]), (tag, s, "matches")
        for u, v in zip(x[5], y[5]):
            assert np.array_equal(u, v), (tag, s, "decisions")


@pytest.mark.parametrize("profile", ["B", "R"])
def test_batch_equals_separate_trackers(ctx, profile):
    print(f"batch of {len(SHAPES)} streams, profile {profile}", flush=True)
    n = len(SHAPES)
    videos = [cu.synth_video(0xC0FFEE40 + i, r, c, 7, i % 2) for i, (r, c) in enumerate(SHAPES)]
    params = [stream_params(i, profile) for i in range(n)]
    batch = cu.Batch(ctx, SHAPES, params)
    solo = [cu.Tracker(ctx, r, c, params[i]) for i, (r, c) in enumerate(SHAPES)]
    for t in solo:
        t.set_graph(False)
    frames = [[ctx.upload(videos[i][f]) for i in range(n)] for f in range(7)]
    for f in range(7):
        batch.step(frames[f])
        for i in range(n):
            solo[i].run(frames[f][i])
        for i in range(n):
            assert_same(state(batch.tracker(i), params[i].nscales), state(solo[i], params[i].nscales), (f, i))
    assert sum(batch.tracker(i).nparticles(0) for i in range(n)) > 100
    batch.close()
    for t in solo:
        t.close()


def test_graph_replay_equals_plain_launches(ctx):
    video = cu.synth_video(0xC0FFEE41, 240, 320, 10, 1)
    p = cu.profile_params("B", 3, 12.0, detector_frequency=2)
    g, e = cu.Tracker(ctx, 240, 320, p), cu.Tracker(ctx, 240, 320, p)
    e.set_graph(False)
    for f in range(10):
        im = ctx.upload(video[f])
        g.run(im)
        e.run(im)
        ctx.sync()
        assert_same(state(g, 3), state(e, 3), f)
        assert g.frame_count == e.frame_count == f + 1


def test_batch_clear_and_tuning(ctx):
    shapes = [(96, 128), (120, 160)]
    params = cu.profile_params("B", 2)
    b = cu.Batch(ctx, shapes, params)
    video = [cu.synth_video(0xC0FFEE42 + i, r, c, 4, 0) for i, (r, c) in enumerate(shapes)]
    frames = [[ctx.upload(video[i][f]) for i in range(2)] for f in range(4)]
    for f in range(3):
        b.step(frames[f])
    ref = state(b.tracker(1), 2)
    b.set_tuning(16, 32)
    assert b.tuning() == (16, 32)
    b.clear()
    assert b.tracker(0).frame_count == 0 and b.tracker(0).nparticles(0) == 0
    for f in range(3):
        b.step(frames[f])
    assert_same(state(b.tracker(1), 2), ref, "after clear + retune")
    with pytest.raises(cu.CuimgError):
        b.tracker(0).run(frames[0][0])  # a stream of a batch steps with the batch
    b.close()


def test_batch_rejects_mixed_profiles(ctx):
    with pytest.raises(cu.CuimgError):
        cu.Batch(ctx, [(96, 128), (96, 128)], [cu.profile_params("B", 2), cu.profile_params("R", 2)])


def test_lk_track_standalone_equals_tracker(ctx):
    """cuimg_hip_lk_track on the tracker's own matcher images and particle
    vector reproduces the tracker's matcher decisions and moves."""
    video = cu.synth_video(0xC0FFEE43, 200, 280, 3, 1)
    L = 3
    p = cu.profile_params("B", L, 12.0, with_merge=0, detector_frequency=100)
    t = cu.Tracker(ctx, 200, 280, p)
    t.set_graph(False)
    t.run(video[0])
    before = [t.particles(s) for s in range(L)]
    t.run(video[1])
    imgs = [t.matcher_images(s) for s in range(L)]
    prev = [ctx.upload(a) for a, _ in imgs]
    cur = [ctx.upload(b) for _, b in imgs]
    for s in range(L):
        parts, acc, tr, dist = cu.lk_track(ctx, prev, cur, s, before[s], ws=7, max_iters=5, eps=0.1, top_n=49,
                                           k=300)
        acc_t, tr_t, dist_t = t.decisions(s)
        assert np.array_equal(acc, acc_t) and np.array_equal(tr, tr_t) and np.array_equal(dist, dist_t), s
        after = t.particles(s)  # no detection ran: same vector, moved in place
        for k in parts.dtype.names:
            assert np.array_equal(parts[k], after[k]), (s, k)


def test_batch_step_from_host_frames(ctx):
    """batch_step from host frames (gray, and 3-channel with channel 0 the
    gray frame) equals batch_step from device frames, bit for bit."""
    shapes = [(96, 128), (120, 160), (72, 96)]
    params = cu.profile_params("B", 2)
    video = [cu.synth_video(0xC0FFEE50 + i, r, c, 5, 1) for i, (r, c) in enumerate(shapes)]
    rng = np.random.default_rng(3)
    bd, bh, bc = (cu.Batch(ctx, shapes, params) for _ in range(3))
    for f in range(5):
        print(f"batch host frame {f}", flush=True)
        dev = [ctx.upload(video[i][f]) for i in range(len(shapes))]
        bd.step(dev)
        bh.step([video[i][f] for i in range(len(shapes))])
        rgb = []
        for i, (r, c) in enumerate(shapes):
            x = rng.integers(0, 256, (r, c, 3), dtype=np.uint8)
            x[:, :, 0] = video[i][f]
            rgb.append(x)
        bc.step(rgb)
        for d in dev:
            d.free()
    for i in range(len(shapes)):
        for s in range(2):
            want = bd.tracker(i).particles(s)
            for b in (bh, bc):
                got = b.tracker(i).particles(s)
                assert got.tobytes() == want.tobytes(), (i, s)
    for b in (bd, bh, bc):
        b.close()
