"""GPU parity of the bc2s path (the reference sample's default tracker,
generic_strategy<bc2s_feature, fast_detector>): integer arithmetic only, so
every comparison with the oracle is bit-exact.

  * the integer 1-2-1 blur, the integer-tap Gaussian, the feature images and
    the contrast on ragged and benchmark sizes (borders included);
  * the descriptor and the gradient-descent matcher (both scales and the
    two-step match) at random points;
  * the frame step in lockstep with the oracle tracker (particles,
    descriptors, decisions, sparse image, saliency, flow, multiscale count,
    matches) on translation / affine / integer-shift videos, 1080p included,
    through run, update_input + step, clear_scale and set_particles;
  * a batch of bc2s streams equal to their own trackers.
"""
import numpy as np
import pytest

import cuimg_amd as cu
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


def shift_video(seed, nr, nc, nframes, dr, dc):
    """A smooth random texture moved by (dr, dc) whole pixels per frame."""
    rng = np.random.default_rng(seed)
    pad = 4 + max(abs(dr), abs(dc)) * nframes
    big = rng.integers(0, 256, (nr + 2 * pad, nc + 2 * pad)).astype(np.float32)
    k = np.array([1, 4, 6, 4, 1], np.float32) / 16
    for ax in (0, 1):
        big = np.apply_along_axis(lambda v: np.convolve(v, k, "same"), ax, big)
    big = ((big - big.min()) / (np.ptp(big) + 1e-9) * 255).astype(np.uint8)
    return np.stack([big[pad - f * dr:pad - f * dr + nr, pad - f * dc:pad - f * dc + nc] for f in range(nframes)])


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (17, 5), (97, 131), (240, 320), (1080, 1920)])
def test_blur121_and_feature_images(ctx, orc, shape):
    nr, nc = shape
    img = np.random.default_rng(nr * 7 + nc).integers(0, 256, shape).astype(np.uint8)
    print(f"bc2s images {shape}", flush=True)
    src = ctx.upload(img)
    s1 = ctx.image(nr, nc, cu.U8, border=3)
    s2 = ctx.image(nr, nc, cu.U8, border=6)
    c32 = ctx.image(nr, nc, cu.U32)
    c8 = ctx.image(nr, nc, cu.U8)
    ctx.bc2s_images(src, s1, s2, c32)
    ref1, ref2 = orc.bc2s_images(img)
    con = orc.contrast(img)
    assert np.array_equal(s1.download(), ref1)
    assert np.array_equal(s2.download(), ref2)
    assert np.array_equal(c32.download(), con)
    # borders clamp-filled
    assert np.array_equal(s1.download_with_border(), np.pad(ref1, 3, mode="edge"))
    assert np.array_equal(s2.download_with_border(), np.pad(ref2, 6, mode="edge"))
    ctx.bc2s_images(src, s1, s2, c8)
    assert np.array_equal(c8.download(), np.minimum(con, 255).astype(np.uint8))
    one = ctx.image(nr, nc, cu.U8)
    ctx.blur121(src, one)
    assert np.array_equal(one.download(), ref1)
    for im in (src, s1, s2, c32, c8, one):
        im.free()


@pytest.mark.parametrize("sigma,hw", [(2.0, 3), (1.0, 1), (1.5, 4), (3.0, 9)])
@pytest.mark.parametrize("shape", [(5, 3), (240, 320), (1080, 1920)])
def test_integer_gaussian(ctx, orc, sigma, hw, shape):
    img = np.random.default_rng(hw).integers(0, 256, shape).astype(np.uint8)
    src, dst = ctx.upload(img), ctx.image(shape[0], shape[1], cu.U8)
    ctx.gaussian_int(src, dst, sigma, hw)
    assert np.array_equal(dst.download(), orc.gaussian_int(img, sigma, hw))
    src.free()
    dst.free()


def test_descriptor_and_descent(ctx, orc):
    v = shift_video(5, 120, 160, 2, 2, -3)
    s1h, s2h = orc.bc2s_images(v[0])
    t1h, t2h = orc.bc2s_images(v[1])
    a = ctx.upload(v[0])
    b = ctx.upload(v[1])
    imgs = []
    for src in (a, b):
        s1 = ctx.image(120, 160, cu.U8, border=3)
        s2 = ctx.image(120, 160, cu.U8, border=6)
        ctx.bc2s_images(src, s1, s2)
        imgs.append((s1, s2))
    rng = np.random.default_rng(0)
    pts = np.stack([rng.integers(0, 120, 400), rng.integers(0, 160, 400)], 1).astype(np.int32)
    feats = ctx.bc2s_features(imgs[0][0], imgs[0][1], pts)
    for i, (r, c) in enumerate(pts):
        assert np.array_equal(feats[i], orc.bc2s_feature(s1h, s2h, r, c)), (r, c)
    # predictions around the true match, a few outside the domain
    preds = pts + np.array([2, -3]) + rng.integers(-3, 4, pts.shape)
    preds[:5] = [[-1, 5], [5, -2], [120, 3], [3, 160], [-4, -4]]
    for scale in (0, 1, 2):
        disp, dist = ctx.bc2s_match(imgs[1][0], imgs[1][1], feats, preds, scale)
        for i in range(len(pts)):
            d, ds = orc.bc2s_match(t1h, t2h, feats[i], preds[i], scale)
            assert tuple(disp[i]) == d and dist[i] == ds, (scale, i, preds[i], disp[i], d, dist[i], ds)
    # interior points of the exact shift land on the shifted position
    disp, _ = ctx.bc2s_match(imgs[1][0], imgs[1][1], feats, pts + np.array([2, -3]), 0)
    inner = (pts[:, 0] > 10) & (pts[:, 0] < 108) & (pts[:, 1] > 10) & (pts[:, 1] < 148)
    assert (np.abs(disp[inner]).sum(1) == 0).mean() > 0.95
    for pair in imgs:
        for im in pair:
            im.free()
    a.free()
    b.free()


def state_equal(g, t, nscales, what):
    for s in range(nscales):
        gp, tp = g.particles(s), t.particles(s)
        assert gp.tobytes() == tp.tobytes(), (what, s, "particles", len(gp), len(tp))
        assert np.array_equal(g.features(s), t.features(s)), (what, s, "features")
        assert np.array_equal(g.sparse(s), t.sparse(s)), (what, s, "sparse")
        assert g.compact_has_run(s) == t.compact_has_run(s), (what, s)
        if t.compact_has_run(s):
            assert np.array_equal(g.matches(s), t.matches(s)), (what, s, "matches")
        gf, gm = g.flow(s)
        tf, tm = t.flow(s)
        assert np.array_equal(gm, tm), (what, s, "multiscale count")
        assert np.array_equal(gf.view(np.uint32), tf.view(np.uint32)), (what, s, "flow")
        a1, a2, ac = g.bc2s_images(s)
        b1, b2, bc = t.bc2s_images(s)
        assert np.array_equal(a1, b1) and np.array_equal(a2, b2), (what, s, "feature images")
        assert np.array_equal(ac, np.minimum(bc, 255)), (what, s, "contrast")
        ga, gt, gd = g.decisions(s)
        ta, tt, td = t.last_decisions(s)
        n = len(ta)
        assert np.array_equal(ga[:n], ta) and np.array_equal(gt[:n], tt) and np.array_equal(gd[:n], td), \
            (what, s, "decisions")


def run_pair(ctx, orc, video, nscales, th, nframes=None, **kw):
    nr, nc = video.shape[1:]
    g = cu.Tracker(ctx, nr, nc, cu.profile_params("C", nscales, th, **kw))
    t = orc.tracker(nr, nc, oracle.profile_params("C", nscales, th, **kw))
    nframes = nframes or len(video)
    tracked = 0
    for f in range(nframes):
        g.run(video[f])
        t.run(video[f])
        state_equal(g, t, nscales, f)
        tracked += int((t.particles(0)["age"] > 1).sum())
    return g, t, tracked


@pytest.mark.parametrize("motion,th,det", [(0, 20.0, 1), (1, 20.0, 10), (1, 8.0, 3)])
def test_bc2s_lockstep_synthetic(ctx, orc, motion, th, det):
    video = cu.synth_video(0xC0FFEE03 + motion, 240, 320, 12, motion)
    print(f"bc2s lockstep motion {motion} th {th} det {det}", flush=True)
    g, t, tracked = run_pair(ctx, orc, video, 3, th, detector_frequency=det)
    print(f"  {tracked} tracked particle-frames", flush=True)
    assert tracked > 200
    g.close()


def test_bc2s_lockstep_integer_shift_known_answer(ctx, orc):
    """An exact (-1, +2) px/frame shift: most tracked particles carry it."""
    video = shift_video(11, 200, 264, 10, -1, 2)
    g, t, tracked = run_pair(ctx, orc, video, 3, 20.0, detector_frequency=4)
    p = g.particles(0)
    live = p[p["age"] > 1]
    assert len(live) > 500
    assert ((live["speed_r"] == -1) & (live["speed_c"] == 2)).mean() > 0.9
    g.close()


@pytest.mark.parametrize("fast_th", [8.0, 20.0])
def test_bc2s_lockstep_1080p(ctx, orc, fast_th):
    """The sample's settings (detector every 10 frames, filtering every
    frame) at 1920x1080 with 4 scales for 20 frames."""
    video = cu.synth_video(0xC0FFEE02, 1080, 1920, 20, 1)
    print(f"bc2s 1080p lockstep th {fast_th}", flush=True)
    g, t, tracked = run_pair(ctx, orc, video, 4, fast_th, detector_frequency=10)
    print(f"  {tracked} tracked particle-frames", flush=True)
    assert tracked > 20_000
    g.close()


def test_bc2s_update_input_step_clear_scale_set_particles(ctx, orc):
    video = cu.synth_video(0xC0FFEE05, 160, 224, 10, 1)
    params_g = cu.profile_params("C", 3, 12.0, detector_frequency=3)
    params_t = oracle.profile_params("C", 3, 12.0, detector_frequency=3)
    g = cu.Tracker(ctx, 160, 224, params_g)
    t = orc.tracker(160, 224, params_t)
    for f in range(10):
        if f == 4:
            g.clear_scale(1)
            t.clear_scale(1)
        if f == 6:
            # particles replaced from outside: descriptors taken from the
            # current feature images
            p = t.particles(0)[::2].copy()
            g.set_particles(0, p)
            t.set_particles(0, p)
            state_equal(g, t, 3, "set_particles")
        if f % 2:
            g.update_input(video[f])
            g.step()
        else:
            g.run(video[f])
        t.update_input(video[f])
        t.step()
        state_equal(g, t, 3, f)
    g.close()


def test_bc2s_batch_streams_equal_own_trackers(ctx):
    shapes = [(120, 160), (97, 131), (240, 320)]
    params = [cu.profile_params("C", n, 15.0, detector_frequency=2) for n in (2, 3, 3)]
    videos = [cu.synth_video(0xC0FFEE20 + i, h, w, 6, 1) for i, (h, w) in enumerate(shapes)]
    batch = cu.Batch(ctx, shapes, params)
    solo = [cu.Tracker(ctx, h, w, params[i]) for i, (h, w) in enumerate(shapes)]
    views = [batch.tracker(i) for i in range(len(shapes))]
    for f in range(6):
        frames = [ctx.upload(videos[i][f]) for i in range(len(shapes))]
        batch.step(frames)
        for i in range(len(shapes)):
            solo[i].run(frames[i])
        for i in range(len(shapes)):
            for s in range(params[i].nscales):
                assert views[i].particles(s).tobytes() == solo[i].particles(s).tobytes(), (f, i, s)
                assert np.array_equal(views[i].features(s), solo[i].features(s)), (f, i, s)
        for im in frames:
            im.free()
    for t in solo:
        t.close()
    batch.close()


def test_bc2s_free_run_matches_committed_fixture(ctx):
    """Ten free-running frames of the 320x240 video with the reference
    sample's detector frequency: every frame equals the committed fixture."""
    import json
    import os

    import golden.make_frame_step_fixture as m
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frame_step_320x240.json")) as f:
        want = json.load(f)["runs"]["C/exact"]
    t = cu.Tracker(ctx, m.NR, m.NC, cu.profile_params("C", m.NSCALES, m.TH, detector_frequency=m.C_DET_FREQ))
    v = m.video()
    for f in range(m.NFRAMES):
        t.run(v[f])
        assert m.state_bc2s(t, m.NSCALES) == want[f], f
    t.close()


def test_cpp_bc2s_sample_facade_matches_c_abi():
    """The reference sample's own strategy and settings through the C++ facade
    (tracker<tracking_strategies::bc2s_fast_gradient_cpu>): trajectories kept
    with sync_attributes, the facade's particles and descriptors equal a raw C
    ABI tracker's; the sample checks it and prints OK."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuimg_amd", "bin",
                       "bc2s_sample")
    if not os.path.exists(exe):
        pytest.skip("the C++ sample is not built")
    print("bc2s_sample 3 20 14 240 320", flush=True)
    r = subprocess.run([exe, "3", "20", "14", "240", "320"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.strip().endswith("OK")

