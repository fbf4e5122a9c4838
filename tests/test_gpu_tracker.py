"""GPU tracker vs the CPU oracle, frame by frame, through the C ABI (the
lockstep comparison and its bars are described in tests/lockstep.py)."""
import numpy as np
import pytest

import cuimg_amd as cu
import oracle

pytestmark = pytest.mark.gpu

from lockstep import TOL_PX, run_lockstep, to_cu  # noqa: F401
from lockstep import compare_frame as lockstep_compare
from lockstep import new_stats as lockstep_new_stats


def assert_bit_exact(st, nframes):
    """With the oracle in the GPU's arithmetic every matcher decision
    (accept flag, translation and distance bits) is equal and every frame
    compares exactly in full (levels, saliency, dense vectors, matches, sparse
    image)."""
    assert st["bit_exact"] == st["decisions"] > 0, st
    assert st["decision_mismatch"] == 0 and st["within"] == st["tracked"], st
    assert st["frames_exact"] == nframes, st


@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("profile", ["B", "R"])
def test_lockstep_config1_size(ctx, orc, profile):
    video = cu.synth_video(0xC0FFEE01, 240, 320, 12, 0)
    p_o = oracle.profile_params(profile, 3)
    p_g = cu.profile_params(profile, 3)
    st = run_lockstep(ctx, orc, video, p_o, p_g, 12)
    print(st)
    assert st["tracked"] > 1000
    assert_bit_exact(st, 12)


def test_lockstep_affine_4_levels(ctx, orc):
    video = cu.synth_video(0xC0FFEE02, 360, 640, 8, 1)
    st = run_lockstep(ctx, orc, video, oracle.profile_params("B", 4), cu.profile_params("B", 4), 8)
    print(st)
    assert st["tracked"] > 10000
    assert_bit_exact(st, 8)


def test_detector_and_merge_frequencies(ctx, orc):
    video = cu.synth_video(0xC0FFEE05, 200, 280, 10, 0)
    kw = dict(detector_frequency=3, filtering_frequency=2)
    st = run_lockstep(ctx, orc, video, oracle.profile_params("B", 3, **kw), cu.profile_params("B", 3, **kw), 10)
    print(st)
    assert_bit_exact(st, 10)


def test_independent_run_statistics(ctx, orc):
    """Both trackers run on their own for 30 frames: the particle populations
    stay statistically identical (counts within 1 %)."""
    video = cu.synth_video(0xC0FFEE01, 240, 320, 30, 0)
    o = orc.tracker(240, 320, oracle.profile_params("B", 3))
    g = cu.Tracker(ctx, 240, 320, cu.profile_params("B", 3))
    for f in range(30):
        o.run(video[f])
        g.run(video[f])
    for s in range(3):
        no, ng = len(o.particles(s)), len(g.particles(s))
        assert abs(no - ng) <= max(2, 0.01 * no), (s, no, ng)


def test_clear_resets(ctx):
    video = cu.synth_video(0xC0FFEE01, 120, 160, 3, 0)
    g = cu.Tracker(ctx, 120, 160, cu.profile_params("B", 2))
    for f in range(3):
        g.run(video[f])
    assert g.nparticles(0) > 0
    g.clear()
    assert g.nparticles(0) == 0 and g.frame_count == 0
    assert (g.sparse(0) == -1).all()
    g.run(video[0])
    assert g.nparticles(0) > 0


def test_bad_arguments_raise(ctx):
    with pytest.raises(cu.CuimgError):
        cu.Tracker(ctx, 100, 100, cu.profile_params("B", 9))
    with pytest.raises(cu.CuimgError):
        cu.Tracker(ctx, 100, 100, cu.profile_params("B", 2, lk_ws=8))
    with pytest.raises(cu.CuimgError):  # the matcher's 32-bit offsets: at most 16384 per side
        cu.Tracker(ctx, 64, 20000, cu.profile_params("B", 2))
    g = cu.Tracker(ctx, 64, 64, cu.profile_params("B", 2))
    with pytest.raises(cu.CuimgError):
        g.run(ctx.image(32, 32, cu.U8))


def test_host_frames_match_device_frames(ctx):
    """run() from host memory (pinned double buffer, upload on the copy stream)
    and from interleaved 3- and 4-channel host frames (channel 0 is the gray
    frame, as get_x) give the same tracker state, bit for bit, as run() on
    device-resident frames."""
    video = cu.synth_video(0xC0FFEE01, 240, 320, 10, 1)
    rng = np.random.default_rng(9)
    params = cu.profile_params("B", 3)
    dev = cu.Tracker(ctx, 240, 320, params)
    host = cu.Tracker(ctx, 240, 320, params)
    rgb = cu.Tracker(ctx, 240, 320, params)
    rgba = cu.Tracker(ctx, 240, 320, params)
    padded = np.zeros((240, 400), np.uint8)  # a host pitch wider than the row
    for f in range(10):
        print(f"host frame {f}", flush=True)
        img = ctx.image(240, 320, cu.U8)
        img.upload(video[f])
        dev.run(img)
        padded[:, :320] = video[f]
        host.run(padded[:, :320])
        c3 = rng.integers(0, 256, (240, 320, 3), dtype=np.uint8)
        c3[:, :, 0] = video[f]
        rgb.run(c3)
        c4 = rng.integers(0, 256, (240, 320, 4), dtype=np.uint8)
        c4[:, :, 0] = video[f]
        rgba.run(c4)
        img.free()
    for s in range(3):
        want = dev.particles(s)
        assert len(want) > 0
        for t in (host, rgb, rgba):
            got = t.particles(s)
            assert len(got) == len(want) and got.tobytes() == want.tobytes(), s
            assert np.array_equal(t.sparse(s), dev.sparse(s))


def test_cpp_sample_facade_matches_c_abi():
    """The C++ facade (the reference sample's call pattern, host frames through
    the staged upload) and a tracker driven through the raw C ABI end in the
    same state; the sample checks it and prints OK."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuimg_amd", "bin",
                       "tracking_sample")
    if not os.path.exists(exe):
        pytest.skip("the C++ sample is not built")
    print("tracking_sample 3 20 8 240 320", flush=True)
    r = subprocess.run([exe, "3", "20", "8", "240", "320"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.strip().endswith("OK")


@pytest.mark.parametrize("fast_th", [8.0, 20.0])
def test_lockstep_1080p_full_size(ctx, orc, fast_th):
    """The headline workload's size (1920x1080, 4 scales, ~100k particles at
    FAST th 8; the reference's default th 20 beside it) for 30 frames in
    lockstep with the oracle: bit-exact like the small sizes."""
    video = cu.synth_video(0xC0FFEE02, 1080, 1920, 30, 1)
    st = run_lockstep(ctx, orc, video, oracle.profile_params("B", 4, fast_th), cu.profile_params("B", 4, fast_th),
                      30)
    print(st)
    assert st["tracked"] > (1_500_000 if fast_th == 8.0 else 300_000)
    assert_bit_exact(st, 30)
    assert st["partial_checked"] >= 0.99 * (st["partial_checked"] + st["partial_skipped"])


@pytest.mark.parametrize("shape,levels", [((251, 333), 3), ((131, 517), 4)])
def test_lockstep_ragged_sizes(ctx, orc, shape, levels):
    """Odd level sizes (floor halving), partial tiles and boxes at every level."""
    nr, nc = shape
    video = cu.synth_video(0xC0FFEE07, nr, nc, 6, 1)
    st = run_lockstep(ctx, orc, video, oracle.profile_params("B", levels, 12.0),
                      cu.profile_params("B", levels, 12.0), 6)
    print(st)
    assert st["tracked"] > 500
    assert_bit_exact(st, 6)


def test_static_video_known_answer(ctx):
    """The oracle's static-video invariants (tests/test_oracle_lk.py) on the GPU
    tracker."""
    import test_oracle_lk as t
    frame = t._texture(5, 160, 224)
    g = cu.Tracker(ctx, 160, 224, cu.profile_params("B", 3, 20.0))
    t._static_invariants(g, 3, 5, frame)


def test_host_frame_views_and_shapes(ctx):
    """Host frames are checked against the tracker's shape before the C call,
    and a flipped view tracks exactly like its contiguous copy."""
    video = cu.synth_video(0xC0FFEE01, 120, 160, 4, 1)
    params = cu.profile_params("B", 2)
    a = cu.Tracker(ctx, 120, 160, params)
    b = cu.Tracker(ctx, 120, 160, params)
    with pytest.raises(ValueError):
        a.run(video[0][:119])
    with pytest.raises(ValueError):
        a.run(np.zeros((120, 160, 2), np.uint8))
    for f in range(4):
        flipped = np.ascontiguousarray(video[f][::-1])
        a.run(flipped[::-1])  # a view with a negative row stride
        b.run(np.ascontiguousarray(video[f]))
    for s in range(2):
        assert a.particles(s).tobytes() == b.particles(s).tobytes()
    a.close()
    b.close()


def test_matcher_settings_are_validated(ctx):
    with pytest.raises(cu.CuimgError):
        cu.Tracker(ctx, 64, 64, cu.profile_params("B", 2, lk_top_n=0))
    with pytest.raises(cu.CuimgError):
        cu.Tracker(ctx, 64, 64, cu.profile_params("B", 2, lk_max_iters=-1))
    # a selection above ws*ws means every pixel, like the standalone matcher
    video = cu.synth_video(0xC0FFEE01, 120, 160, 3, 0)
    a = cu.Tracker(ctx, 120, 160, cu.profile_params("B", 2, lk_top_n=1000))
    b = cu.Tracker(ctx, 120, 160, cu.profile_params("B", 2))
    for f in range(3):
        a.run(video[f])
        b.run(video[f])
    assert a.particles(0).tobytes() == b.particles(0).tobytes()


def test_context_close_destroys_live_trackers():
    """Closing a context destroys the trackers and batches made from it; their
    Python handles are then only forgotten."""
    c = cu.Context(0)
    video = cu.synth_video(0xC0FFEE01, 96, 128, 2, 0)
    t = cu.Tracker(c, 96, 128, cu.profile_params("B", 2))
    bt = cu.Batch(c, [(96, 128), (64, 96)], cu.profile_params("B", 2))
    t.run(video[0])
    c.close()
    t.close()
    bt.close()


def test_translation_cadence_known_answer(ctx):
    """The oracle's frame-step known answer (tests/frame_step_kat.py) on the GPU
    tracker: detector / compaction every 3rd frame, merge every 2nd, ages,
    exact shifts, birth pixels and the sparse image's last writer."""
    import frame_step_kat as kat
    p = cu.profile_params("B", 2, 20.0, **kat.CADENCE)
    seen = kat.check_translation_cadence(cu.Tracker(ctx, kat.NR, kat.NC, p))
    print(seen)
    assert seen["interior"] > 20000 and seen["accepted"] > 20000 and seen["merged"] > 0
    assert seen["mistracked"] <= 0.002 * seen["interior"]


def test_crafted_merges_known_answer(ctx, orc):
    import frame_step_kat as kat
    p = cu.profile_params("B", 1, 20.0, detector_frequency=1000, filtering_frequency=1)
    kat.check_crafted_merges(lambda: cu.Tracker(ctx, kat.NR, kat.NC, p), orc, cu.PARTICLE_DTYPE)


def test_lockstep_raster_order_statistic(ctx, orc):
    """With the oracle's window sums in the reference's raster order the GPU's
    decisions differ only by FP32 reassociation: the same 0.01 px bar."""
    video = cu.synth_video(0xC0FFEE01, 240, 320, 12, 0)
    st = run_lockstep(ctx, orc, video, oracle.profile_params("B", 3), cu.profile_params("B", 3), 12,
                      sum_order="raster")
    print(st)
    assert st["tracked"] > 1000
    assert st["within"] >= 0.999 * st["tracked"]
    assert st["decision_mismatch"] <= 0.001 * st["decisions"]


@pytest.mark.parametrize("profile", ["B", "R"])
def test_free_run_matches_committed_fixture(ctx, profile):
    """Ten free-running frames of the 320x240 video (no state handed over from
    the oracle): every frame's dense vector, sparse image, saliency and
    matches equal the committed fixture of the oracle in the GPU's arithmetic."""
    import json
    import os

    import golden.make_frame_step_fixture as m
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frame_step_320x240.json")) as f:
        want = json.load(f)["runs"][f"{profile}/gpu"]
    t = cu.Tracker(ctx, m.NR, m.NC, cu.profile_params(profile, m.NSCALES, m.TH))
    v = m.video()
    for f in range(m.NFRAMES):
        t.run(v[f])
        assert m.state(t, m.NSCALES) == want[f], (profile, f)
    t.close()


def test_per_scale_box_size_and_particle_domain(ctx, orc):
    """fast_detector::set_box_size and pyrlk_cpu::set_particle_domain per scale
    (changed before the first frame and again mid-run) in lockstep with the
    oracle: bit-exact."""
    video = cu.synth_video(0xC0FFEE01, 240, 320, 10, 1)
    p_o = oracle.profile_params("B", 3, 12.0, sum_lanes=oracle.gpu_sum_lanes(7))
    o = orc.tracker(240, 320, p_o)
    g = cu.Tracker(ctx, 240, 320, cu.profile_params("B", 3, 12.0))
    st = lockstep_new_stats()
    settings = {0: [(0, 3, (20, 30, 200, 280)), (1, 5, (3, 3, 116, 156))],
                5: [(0, 2, (3, 3, 236, 316)), (2, 4, (10, 10, 50, 70))]}
    for f in range(10):
        for s, box, dom in settings.get(f, []):
            o.set_scale(s, box, dom)
            sp = g.scale_params(s)
            sp.box_size = box
            sp.pd_r1, sp.pd_c1, sp.pd_r2, sp.pd_c2 = dom
            g.set_scale_params(s, sp)
        pre = [o.particles(s) for s in range(3)]
        for s in range(3):
            g.set_particles(s, to_cu(pre[s]))
        o.run(video[f])
        g.run(video[f])
        lockstep_compare(o, g, 3, st, pre)
    print(st)
    assert st["tracked"] > 500
    assert_bit_exact(st, 10)
    # particles outside a scale's domain are removed
    p = g.particles(2)
    live = p[p["age"] > 1]
    assert len(live) and (live["pos_r"] >= 10).all() and (live["pos_r"] < 51).all()


def test_update_input_then_step_equals_run(ctx, orc):
    """tracker::update_input + tracker::run() (and run() with no new input,
    which steps again on the last input), plus the strategy's clear() of one
    scale, against the oracle doing the same: bit-exact."""
    video = cu.synth_video(0xC0FFEE03, 240, 320, 8, 1)
    p_o = oracle.profile_params("B", 3, 12.0, sum_lanes=oracle.gpu_sum_lanes(7))
    o = orc.tracker(240, 320, p_o)
    g = cu.Tracker(ctx, 240, 320, cu.profile_params("B", 3, 12.0))
    dev = ctx.image(240, 320, cu.U8)
    for f in range(8):
        if f == 4:  # run() on the previous input again
            o.step()
            g.step()
        else:
            o.update_input(video[f])
            if f % 2:
                dev.upload(video[f])
                g.update_input(dev)
            else:
                g.update_input(video[f])
            o.step()
            g.step()
        if f == 5:
            o.clear_scale(1)
            g.clear_scale(1)
        for s in range(3):
            assert np.array_equal(o.level(s), g.level(s)), (f, s)
            po, pg = o.particles(s), g.particles(s)
            assert len(po) == len(pg), (f, s)
            for k in ("speed_r", "speed_c", "pos_r", "pos_c", "age"):
                assert np.array_equal(po[k], pg[k]), (f, s, k)
            assert np.array_equal(o.sparse(s), g.sparse(s)), (f, s)
    assert sum(len(g.particles(s)) for s in range(3)) > 300


def _near_power_of_two(nr, nc, rng):
    """Particles whose row or column lies within 4.6 px below (or 2.5 above)
    2^k, in pairs one ulp apart (both parities of the last mantissa bit), with
    three speeds; the other coordinate uniform inside the frame."""
    out = []
    for axis, n, m in ((0, nr, nc), (1, nc, nr)):
        k = 3
        while (1 << k) + 12 < n:
            for d in np.linspace(-2.5, 4.6, 15):
                x = np.float32((1 << k) - d)
                for xv in (x, np.nextafter(x, np.float32(0))):
                    for sp in (0.0, 0.37, -1.45):
                        o = np.float32(rng.uniform(16, m - 16))
                        out.append((xv, o, sp) if axis == 0 else (o, xv, sp))
            k += 1
    p = np.zeros(len(out), oracle.PARTICLE_DTYPE)
    for i, (r, c, sp) in enumerate(out):
        p[i]["pos_r"], p[i]["pos_c"] = r, c
        p[i]["speed_r"] = p[i]["speed_c"] = sp
        p[i]["age"] = 2
    return p


def test_matcher_windows_across_powers_of_two(ctx, orc):
    """Windows whose coordinate + half-width crosses a power of two take the
    matcher's per-sample path unless the coordinate's last mantissa bit is zero
    (then every integer offset is exact): particles seeded on both sides of
    that rule at every scale of a 1080 x 1920 frame decide bit for bit like
    the oracle."""
    nr, nc, L = 1080, 1920, 4
    video = cu.synth_video(0xC0FFEE05, nr, nc, 2, 1)
    o = orc.tracker(nr, nc, oracle.profile_params("B", L, 20.0, sum_lanes=oracle.gpu_sum_lanes(7)))
    g = cu.Tracker(ctx, nr, nc, cu.profile_params("B", L, 20.0))
    o.run(video[0])
    g.run(video[0])
    rng = np.random.default_rng(5)
    pre = [_near_power_of_two(nr >> s, nc >> s, rng) for s in range(L)]
    crossing = {0: 0, 1: 0}
    for s in range(L):
        for f in ("pos_r", "pos_c"):
            x = pre[s][f]
            top = 2.0 ** np.floor(np.log2(x.astype(np.float64)) + 1)
            cross = x.astype(np.float64) + 3 >= top
            lsb = x.view(np.uint32) & 1
            crossing[0] += int((cross & (lsb == 0)).sum())
            crossing[1] += int((cross & (lsb == 1)).sum())
        o.set_particles(s, pre[s])
        g.set_particles(s, to_cu(pre[s]))
    assert min(crossing.values()) > 100, crossing
    o.run(video[1])
    g.run(video[1])
    st = lockstep_new_stats()
    lockstep_compare(o, g, L, st, pre)
    print(st)
    assert st["tracked"] > 1000
    assert_bit_exact(st, 1)
    g.close()
