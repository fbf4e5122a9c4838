"""Benchmark of the tracking hot path on one MI355X.

Workload (default): the plain method on a 1920x1080 uint8 synthetic video with
per-frame random affine motion, 4 pyramid scales, sigma=1 / 7-tap FP32 blur,
7x7 LK windows with <= 5 iterations, FAST threshold 8, which keeps about 100k
particles per frame on this video.  One step = one tracker frame step
(pyramid, blur, matcher on every scale, detector mask, saliency, box
selection, insertion, compaction, merge) on one frame.  All frames are
generated on the host and uploaded before timing, so the timed region starts
with its inputs resident in HBM.  The same video is also run at the
reference's default threshold 20 (about 14k particles) and reported beside it.

value = particle-updates/s: particles (age > 0) advanced through the matcher
at their own scale (each one includes its coarser ancestor solves), summed
over scales and timed frames, divided by the timed wall time.

Printed: one JSON line (rank 0).  Extra keys carry frames/s, per-kernel
timings, the threshold-20 run, the Gaussian / pyramid roofline microbench and
the CPU baseline, and the bc2s strategy (the reference sample's default
tracker) on the same video with the sample's settings.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "particle-updates/s & frames/s @1080p 4-lvl 100k pts; % HBM roofline"
HBM_PEAK_GBS = 8000.0
SEED = 0xC0FFEE02
NR, NC, NSCALES = 1080, 1920, 4
FAST_TH = 8.0          # about 100k particles on the synthetic video
FAST_TH_SPARSE = 20.0  # the reference's default threshold: about 14k particles


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lk_algorithmic_bytes(level_shapes, nparts):
    """Compulsory bytes of one matcher launch (all scales at once): the
    previous and current FP32 frames of every scale read once, plus each
    particle read and written (20 B each way)."""
    img = sum(2 * 4 * r * c for (r, c) in level_shapes)
    return img + 40 * sum(nparts)


KERNELS = ("pyramid_build", "blur_gradient", "fills", "lk_track", "particle_mask", "fast_saliency",
           "detect_count", "scatter_insert", "merge_scan", "merge_round1", "merge_resolve")
BC2S_KERNELS = ("pyramid_build", "bc2s_images", "fast_saliency", "fills", "bc2s_match", "flow_stats", "flow_fusion",
                "particle_mask", "detect_count", "scatter_insert", "filter", "merge_scan", "merge_round1",
                "merge_resolve")
BC2S_DET_FREQ = 10  # the reference sample's detector frequency


def upload_video(cu, ctx, nframes):
    t0 = time.time()
    video = cu.synth_video(SEED, NR, NC, nframes, 1)
    log(f"generated {nframes} frames in {time.time() - t0:.1f}s")
    frames = []
    for f in range(nframes):
        im = ctx.image(NR, NC, cu.U8)
        im.upload(video[f])
        frames.append(im)
    return video, frames


def tracker_bench(cu, ctx, frames, fast_th, warmup, steps, prof_steps, pct_steps=0, profile="B", det_freq=1,
                  kernel_names=KERNELS):
    """warmup untimed frames, `steps` timed frame steps replayed from the
    captured graph, `pct_steps` more graph steps each between two HIP events
    (median and p10 / p90 of the step time), then `prof_steps` more steps with
    plain launches and HIP events around every kernel (the per-kernel
    breakdown)."""
    params = cu.profile_params(profile, NSCALES, fast_th, detector_frequency=det_freq)
    trk = cu.Tracker(ctx, NR, NC, params)
    shapes = [(NR >> s, NC >> s) for s in range(NSCALES)]
    for f in range(warmup):
        trk.run(frames[f])
    ctx.sync()
    live0 = [trk.work(s)[0] for s in range(NSCALES)]
    t0 = time.perf_counter()
    for f in range(warmup, warmup + steps):
        trk.run(frames[f])
    ctx.sync()
    dt = time.perf_counter() - t0
    live1 = [trk.work(s)[0] for s in range(NSCALES)]
    per_scale = [b - a for a, b in zip(live0, live1)]
    nparts = [trk.nparticles(s) for s in range(NSCALES)]
    step_ms = []
    for f in range(warmup + steps, warmup + steps + pct_steps):
        ctx.timer_start()
        trk.run(frames[f])
        step_ms.append(ctx.timer_stop())
    live_p = [trk.work(s)[0] for s in range(NSCALES)]
    kernels, prof = {}, None
    if prof_steps:
        trk.set_graph(False)
        ctx.profile(True)
        ctx.profile_reset()
        t1 = time.perf_counter()
        for f in range(warmup + steps + pct_steps, warmup + steps + pct_steps + prof_steps):
            trk.run(frames[f])
        ctx.sync()
        pdt = time.perf_counter() - t1
        live2 = [trk.work(s)[0] for s in range(NSCALES)]
        for name in kernel_names:
            n, ms = ctx.profile_query(name)
            if n:
                kernels[name] = {"launches": n, "total_ms": round(ms, 4), "avg_us": round(ms / n * 1000, 3)}
        ctx.profile(False)
        prof = dict(frames=prof_steps, seconds=pdt, per_scale=[b - a for a, b in zip(live_p, live2)])
    trk.close()
    return dict(seconds=dt, frames=steps, updates=sum(per_scale), per_scale=per_scale, nparts=nparts,
                kernels=kernels, shapes=shapes, fast_th=fast_th, prof=prof, step_ms=step_ms)


def host_frames_bench(cu, ctx, video, fast_th, warmup=20, steps=100):
    """The same tracker fed from host memory (numpy frames): each call stages
    the frame in a pinned buffer and the upload overlaps the previous step.
    The PCIe-inclusive rate, reported beside `value` (never as it)."""
    trk = cu.Tracker(ctx, NR, NC, cu.profile_params("B", NSCALES, fast_th))
    for f in range(warmup):
        trk.run(video[f])
    ctx.sync()
    live0 = sum(trk.work(s)[0] for s in range(NSCALES))
    t0 = time.perf_counter()
    for f in range(warmup, warmup + steps):
        trk.run(video[f])
    ctx.sync()
    dt = time.perf_counter() - t0
    live1 = sum(trk.work(s)[0] for s in range(NSCALES))
    trk.close()
    return {"frames_per_s": round(steps / dt, 2), "particle_updates_per_s": round((live1 - live0) / dt, 1),
            "ms_per_step": round(dt / steps * 1000, 4),
            "workload": f"1920x1080 u8 frames in host memory, {steps} steps after {warmup}, FAST th {fast_th:g}"}


def percentiles(xs):
    if not xs:
        return None
    p10, p50, p90 = np.percentile(np.asarray(xs, np.float64), [10, 50, 90])
    return {"n": len(xs), "p10": round(float(p10), 4), "median": round(float(p50), 4), "p90": round(float(p90), 4)}


def summary(tr):
    # level solves: a particle at scale s is matched at every level from the
    # coarsest (NSCALES - 1) down to s
    solves = sum(n * (NSCALES - s) for s, n in enumerate(tr["per_scale"]))
    return {"fast_th": tr["fast_th"], "particle_updates_per_s": round(tr["updates"] / tr["seconds"], 1),
            "level_solves_per_s": round(solves / tr["seconds"], 1),
            "step_ms_events": percentiles(tr["step_ms"]),
            "frames_per_s": round(tr["frames"] / tr["seconds"], 2),
            "ms_per_step": round(tr["seconds"] / tr["frames"] * 1000, 4),
            "updates_per_frame": round(tr["updates"] / tr["frames"], 1),
            "particles_per_scale": tr["nparts"], "kernels": tr["kernels"]}


BATCH_STREAMS, BATCH_FRAMES, BATCH_SEED = 64, 8, 0xC0FFEE04


def batch_workload(cu):
    """64 streams, width U[320,1920], height round(9w/16), 3 scales when
    min(h, w) < 360 else 4; each stream gets U[1k,150k] particles by seeded
    placement after its first frame (detector threshold 256, above any u8
    saliency, so no point is added: the count is the placement's), clamped
    to the container capacity of one particle per 4 pixels; scale s gets
    n / 4^s."""
    rng = np.random.default_rng(BATCH_SEED)
    shapes, params, parts = [], [], []
    for i in range(BATCH_STREAMS):
        w = int(rng.integers(320, 1921))
        h = int(round(9 * w / 16))
        L = 3 if min(h, w) < 360 else 4
        n = min(int(rng.integers(1000, 150001)), h * w // 4)
        shapes.append((h, w))
        params.append(cu.profile_params("B", L, 256.0))
        lv = []
        for s in range(L):
            hs, ws = h >> s, w >> s
            m = min(max(n >> (2 * s), 1), hs * ws // 4)
            p = np.zeros(m, cu.PARTICLE_DTYPE)
            p["pos_r"] = rng.integers(3, hs - 3, m) + rng.random(m).astype(np.float32) * np.float32(0.999)
            p["pos_c"] = rng.integers(3, ws - 3, m) + rng.random(m).astype(np.float32) * np.float32(0.999)
            p["age"] = 1
            lv.append(p)
        parts.append(lv)
    return shapes, params, parts


def batch_bench(cu, ctx, tuning=None, sequential=True):
    """One frame step of 64 streams as one graph replay of the batched kernels,
    against the same 64 streams stepped one tracker after the other, with
    plain launches (the baseline of the speedup) and with each tracker
    replaying its own captured graph; 7 timed steps (frames 1..7) after the
    seeded placement.  `tuning` = (lk_ppw, blur_th) skips the autotuner."""
    shapes, params, parts = batch_workload(cu)
    t0 = time.time()
    frames = [[None] * BATCH_STREAMS for _ in range(BATCH_FRAMES)]
    for i, (h, w) in enumerate(shapes):
        v = cu.synth_video(BATCH_SEED + 1 + i, h, w, BATCH_FRAMES, 1)
        for f in range(BATCH_FRAMES):
            frames[f][i] = ctx.upload(v[f])
    log(f"batch: {BATCH_STREAMS} streams, {sum(h * w for h, w in shapes) / 1e6:.1f} Mpx per frame, "
        f"{sum(len(p[0]) for p in parts)} particles at scale 0, inputs ready in {time.time() - t0:.1f}s")

    def place(trackers):
        for i, t in enumerate(trackers):
            for s, p in enumerate(parts[i]):
                t.set_particles(s, p)
        ctx.sync()

    def run_batch(b):
        b.step(frames[0])
        place([b.tracker(i) for i in range(BATCH_STREAMS)])
        t1 = time.perf_counter()
        for f in range(1, BATCH_FRAMES):
            b.step(frames[f])
        ctx.sync()
        return time.perf_counter() - t1

    batch = cu.Batch(ctx, shapes, params)
    if tuning is None:
        log("batch: autotune")
        best, tune = batch.autotune(run_batch)
    else:
        batch.set_tuning(*tuning)
        tune = {}
    tb = []
    for k in range(3):
        log(f"batch: batched run {k + 1} of 3")
        batch.clear()
        tb.append(run_batch(batch))
    live = sum(batch.tracker(i).work(s)[0] for i in range(BATCH_STREAMS) for s in range(params[i].nscales))

    solo = [cu.Tracker(ctx, h, w, params[i]) for i, (h, w) in enumerate(shapes)]

    def run_solo(graph):
        for i, t in enumerate(solo):
            t.set_graph(graph)
            t.clear()
            t.run(frames[0][i])
        place(solo)
        t1 = time.perf_counter()
        for f in range(1, BATCH_FRAMES):
            for i, t in enumerate(solo):
                t.run(frames[f][i])
        ctx.sync()
        return time.perf_counter() - t1

    tg = ts = float("nan")
    if sequential:
        log("batch: sequential trackers, graph replay")
        run_solo(True)  # warm-up and capture
        tg = min(run_solo(True) for _ in range(2))
        log("batch: sequential trackers, plain launches")
        run_solo(False)  # warm-up
        ts = min(run_solo(False) for _ in range(2))
    same = 0
    for i in range(BATCH_STREAMS if sequential else 0):
        ok = True
        for s in range(params[i].nscales):
            a, b = batch.tracker(i).particles(s), solo[i].particles(s)
            ok = ok and len(a) == len(b) and all(np.array_equal(a[k], b[k]) for k in a.dtype.names)
        same += ok
    steps = BATCH_FRAMES - 1
    t_b = min(tb)
    out = {"workload": f"{BATCH_STREAMS} streams, w~U[320,1920], h=round(9w/16), 3-4 scales, profile B, "
                       f"{steps} timed frame steps after seeded placement of U[1k,150k] particles",
           "pixels_per_frame": sum(h * w for h, w in shapes),
           "particles_scale0": sum(len(p[0]) for p in parts),
           "batch_ms_per_step": round(t_b / steps * 1000, 3),
           "sequential_ms_per_step": round(ts / steps * 1000, 3) if sequential else None,
           "speedup": round(ts / t_b, 2) if sequential else None,
           "sequential_graph_ms_per_step": round(tg / steps * 1000, 3) if sequential else None,
           "speedup_vs_graph_replay": round(tg / t_b, 2) if sequential else None,
           "stream_frames_per_s": round(BATCH_STREAMS * steps / t_b, 1),
           "particle_updates_per_s": round(live / tb[-1], 1),
           "streams_identical_to_own_tracker": f"{same}/{BATCH_STREAMS}",
           "tuning": dict(zip(("lk_particles_per_wave", "blur_th"), batch.tuning())),
           "tuning_ms_per_step": {f"{c[0]}/{c[1]}": round(v / steps * 1000, 3) for c, v in tune.items()}}
    for t in solo:
        t.close()
    batch.close()
    for row in frames:
        for im in row:
            im.free()
    return out


ADJ_SEED, ADJ_NR, ADJ_NC, ADJ_N = 0xC0FFEE05, 720, 1280, 50000


def adjoint_bench(cu, ctx, reps=20):
    """Loss + gradients (dL/dd and dL/dI1) of the blurred photometric LK loss:
    1280x720 FP32 frames of the synthetic video, 50k particles uniform over
    the interior, d ~ U[-2, 2]^2, sigma 1 / 7 taps, 7x7 windows; device
    buffers allocated once, span-timed back-to-back calls."""
    import ctypes
    rng = np.random.default_rng(ADJ_SEED)
    v = cu.synth_video(ADJ_SEED, ADJ_NR, ADJ_NC, 2, 1).astype(np.float32)
    I0, I1 = ctx.upload(v[0]), ctx.upload(v[1])
    pos = np.stack([rng.uniform(4, ADJ_NR - 5, ADJ_N), rng.uniform(4, ADJ_NC - 5, ADJ_N)], 1).astype(np.float32)
    d = rng.uniform(-2, 2, (ADJ_N, 2)).astype(np.float32)
    bp, bd = cu.DeviceBuffer.from_array(ctx, pos), cu.DeviceBuffer.from_array(ctx, d)
    tot, lp, gd = cu.DeviceBuffer(ctx, 8), cu.DeviceBuffer(ctx, 4 * ADJ_N), cu.DeviceBuffer(ctx, 8 * ADJ_N)
    scratch = cu.DeviceBuffer(ctx, 8 * ADJ_NR * ADJ_NC)
    gI = ctx.image(ADJ_NR, ADJ_NC, cu.F32)
    L = ctx.L

    def call(_):
        ctx.check(L.cuimg_hip_lk_loss_adjoint(ctx.h, ctypes.byref(I0.d), ctypes.byref(I1.d), bp.p, bd.p, ADJ_N, 1, 3,
                                              7, tot.p, lp.p, gd.p, ctypes.byref(gI.d), scratch.p,
                                              2 * ADJ_NR * ADJ_NC), "lk_loss_adjoint")

    us = span_us(ctx, call, reps)
    out = {"workload": f"{ADJ_NC}x{ADJ_NR} f32, {ADJ_N} particles, d~U[-2,2]^2, sigma 1 / 7 taps, 7x7 windows, "
                       "loss + dL/dd + dL/dI1", "avg_us": round(us, 2),
           "particles_per_s": round(ADJ_N / (us * 1e-6), 1)}
    for b in (bp, bd, tot, lp, gd, scratch):
        b.free()
    for im in (I0, I1, gI):
        im.free()
    return out


def span_us(ctx, launch, reps):
    """Average microseconds per launch of `reps` back-to-back launches, timed by
    two events on the context stream (after one untimed pass over the 4 buffer
    sets)."""
    for i in range(4):
        launch(i)
    ctx.sync()
    ctx.timer_start()
    for i in range(reps):
        launch(i)
    return ctx.timer_stop() * 1000 / reps


def launch_us(ctx, launch, reps):
    """Microseconds of each of `reps` launches, each between two events (the
    spread; span_us gives the back-to-back rate)."""
    out = []
    for i in range(reps):
        ctx.timer_start()
        launch(i)
        out.append(ctx.timer_stop() * 1000)
    return out


def gaussian_microbench(cu, ctx, reps=40):
    """Separable Gaussian on f32 4096x4096 and f32x4 3840x2160 (sigma 1, 2, 4;
    2*ceil(3 sigma)+1 taps), 4 distinct buffer pairs rotated so the working set
    (>= 512 MiB) exceeds the 256 MB Infinity Cache; and the 5-level 2x pyramid
    of the 4096x4096 image.  First, a 256 MiB device-to-device copy (the
    runtime's copy kernel) over 4 rotated buffer pairs: the achievable
    read+write rate, and the reference reading for the HBM traffic counters."""
    rng = np.random.default_rng(3)
    out = []
    nb = 256 << 20
    pairs = [(cu.DeviceBuffer(ctx, nb), cu.DeviceBuffer(ctx, nb)) for _ in range(4)]
    us = span_us(ctx, lambda i: pairs[i % 4][1].copy_from(pairs[i % 4][0]), reps)
    gbs = 2 * nb / us / 1e3
    out.append({"kernel": "copy_d2d", "image": "256 MiB", "avg_us": round(us, 2), "bytes": 2 * nb,
                "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
    for a, b in pairs:
        a.free()
        b.free()
    for kind, shape in (("f32", (4096, 4096)), ("f32x4", (2160, 3840))):
        if kind == "f32":
            img = rng.uniform(0, 255, shape).astype(np.float32)
            dt, es = cu.F32, 4
        else:
            img = rng.uniform(0, 255, shape + (4,)).astype(np.float32)
            dt, es = cu.F32X4, 16
        ins = [ctx.upload(img) for _ in range(4)]
        outs = [ctx.image(shape[0], shape[1], dt) for _ in range(4)]
        nbytes = shape[0] * shape[1] * es * 2
        for sigma in (1, 2, 4):
            log(f"gaussian {kind} sigma {sigma}")
            K = 3 * sigma
            run = lambda i: ctx.gaussian(ins[i % 4], outs[i % 4], sigma, K)  # noqa: E731
            us = span_us(ctx, run, reps)
            gbs = nbytes / us / 1e3
            out.append({"kernel": "gaussian_sep", "image": f"{kind} {shape[1]}x{shape[0]}", "sigma": sigma,
                        "taps": 2 * K + 1, "avg_us": round(us, 2), "bytes": nbytes, "GBps": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4), "launch_us": percentiles(launch_us(ctx, run, reps))})
            if kind == "f32" and sigma == 1:
                # one buffer pair (128 MiB) stays in the 256 MB Infinity Cache
                us = span_us(ctx, lambda i: ctx.gaussian(ins[0], outs[0], sigma, K), reps)
                out.append({"kernel": "gaussian_sep", "image": f"{kind} {shape[1]}x{shape[0]}, one buffer pair "
                            "(hot cache, not the roofline figure)", "sigma": sigma, "taps": 2 * K + 1,
                            "avg_us": round(us, 2), "bytes": nbytes, "GBps": round(nbytes / us / 1e3, 1)})
        for im in ins + outs:
            im.free()
    # 5-level pyramid, in place on level 0
    img = rng.uniform(0, 255, (4096, 4096)).astype(np.float32)
    pyrs = []
    for _ in range(4):
        lv = [ctx.upload(img)]
        for l in range(1, 5):
            r, c = lv[-1].shape
            lv.append(ctx.image(r // 2, c // 2, cu.F32))
        pyrs.append(lv)
    nbytes = 4 * 4096 * 4096 * (1 + 1 / 4 + 1 / 16 + 1 / 64 + 1 / 256)
    run = lambda i: ctx.pyramid(pyrs[i % 4][0], pyrs[i % 4])  # noqa: E731
    us = span_us(ctx, run, reps)
    gbs = nbytes / us / 1e3
    out.append({"kernel": "pyramid_build", "image": "f32 4096x4096, 5 levels", "avg_us": round(us, 2),
                "bytes": int(nbytes), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                "launch_us": percentiles(launch_us(ctx, run, reps))})
    # blur-then-subsample level (sigma 1, 13 taps) of the same image
    decs = [ctx.image(2048, 2048, cu.F32) for _ in range(4)]
    nbytes = 4 * 4096 * 4096 + 4 * 2048 * 2048
    us = span_us(ctx, lambda i: ctx.pyr_down(pyrs[i % 4][0], decs[i % 4], cu.PYR_GAUSS_DECIMATE), reps)
    gbs = nbytes / us / 1e3
    out.append({"kernel": "pyr_gauss_decimate", "image": "f32 4096x4096 -> 2048x2048", "avg_us": round(us, 2),
                "bytes": int(nbytes), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
    for lv in pyrs + [decs]:
        for im in lv:
            im.free()
    return out


def cpu_baseline(video, budget_s, fast_th):
    """The CPU oracle (a restatement of the reference algorithms, OpenMP over
    particles and rows like the reference) on the first frames of the same
    video and threshold, until about `budget_s` seconds of CPU work."""
    import oracle
    orc = oracle.Oracle()
    nr, nc = video.shape[1:]
    t = orc.tracker(nr, nc, oracle.profile_params("B", NSCALES, fast_th))
    t.run(video[0])
    frames, updates, spent = 0, 0, 0.0
    while spent < budget_s and frames + 1 < len(video):
        live = sum(int((t.particles(s)["age"] > 0).sum()) for s in range(NSCALES))
        t0 = time.perf_counter()
        t.run(video[frames + 1])
        spent += time.perf_counter() - t0
        frames += 1
        updates += live
    threads = orc.max_threads()
    # the same frames with one thread, on a fresh oracle tracker
    orc.set_threads(1)
    t1 = orc.tracker(nr, nc, oracle.profile_params("B", NSCALES, fast_th))
    t1.run(video[0])
    f1, u1, s1 = 0, 0, 0.0
    while s1 < budget_s / 3 and f1 + 1 < len(video):
        live = sum(int((t1.particles(s)["age"] > 0).sum()) for s in range(NSCALES))
        t0 = time.perf_counter()
        t1.run(video[f1 + 1])
        s1 += time.perf_counter() - t0
        f1 += 1
        u1 += live
    orc.set_threads(threads)
    return {"value": round(updates / spent, 1), "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle, frames 1..{frames} of the same 1920x1080 video at FAST th {fast_th:g} "
                      f"({spent:.1f} s, {updates} particle updates, {frames / spent:.2f} frames/s)",
            "host": {"cpu_model": cpu_model(), "logical_cpus": os.cpu_count(), "omp_threads": threads},
            "one_thread": {"value": round(u1 / s1, 1), "sample": f"frames 1..{f1}, {s1:.1f} s"}}


def bc2s_cpu_baseline(video, budget_s, fast_th):
    """The CPU oracle's bc2s strategy (serial loops) on the first frames of the
    same video with the reference sample's settings."""
    import oracle
    orc = oracle.Oracle()
    nr, nc = video.shape[1:]
    t = orc.tracker(nr, nc, oracle.profile_params("C", NSCALES, fast_th, detector_frequency=BC2S_DET_FREQ))
    t.run(video[0])
    frames, updates, spent = 0, 0, 0.0
    while spent < budget_s and frames + 1 < len(video):
        live = sum(int((t.particles(s)["age"] > 0).sum()) for s in range(NSCALES))
        t0 = time.perf_counter()
        t.run(video[frames + 1])
        spent += time.perf_counter() - t0
        frames += 1
        updates += live
    return {"value": round(updates / spent, 1), "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "frames_per_s": round(frames / spent, 2),
            "sample": f"CPU oracle bc2s strategy, frames 1..{frames} of the same translation video at FAST th {fast_th:g}, "
                      f"detector every {BC2S_DET_FREQ} frames ({spent:.1f} s, {updates} particle updates)"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# HBM traffic of the matcher launch, measured by rocprofv3 counter runs of
# this benchmark's workload (FETCH_SIZE, WRITE_SIZE, scaled by the readings of
# the 256 MiB copy of the microbenchmark); the summary names the command.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "lk_track_1080p_th8_hbm_traffic.txt")


def measured_traffic(path=TRAFFIC_PROFILE):
    """Bytes per launch from the `traffic_bytes_per_launch = N` line of the
    counter summary, None when there is none."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch"):
                    return int(line.split("=")[1].strip())
    except OSError:
        pass
    return None


# VALU wave instructions of the matcher launch, from a rocprofv3 SQ_INSTS_VALU
# run of this benchmark's workload; the summary names the command.
VALU_PROFILE = os.path.join(ROOT, "profiles", "lk_track_1080p_th8_valu.txt")
# one wave64 VALU instruction per SIMD every 4 cycles: 256 CUs x 4 SIMDs x 2.4 GHz / 4
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


# FP32 peak of MI355X (dense, non-packed), for the matcher's flop fraction
FP32_PEAK_TFLOPS = 157.3
# flops of one 7x7 level solve with at most 5 Gauss-Newton steps, counted as
# 49 x (10 Sobel + 8 bilinear + 6 outer product) + 5 x 49 x (8 + 1 + 4) + 10
LK_FLOPS_PER_SOLVE = 49 * (10 + 8 + 6) + 5 * 49 * (8 + 1 + 4) + 10


def measured_valu(path=VALU_PROFILE):
    """VALU wave instructions per level solve from the `valu_instr_per_level_solve
    = X` line of the counter summary (a separate rocprofv3 run of the same
    workload), None when there is none."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("valu_instr_per_level_solve"):
                    return float(line.split("=")[1].strip())
    except OSError:
        pass
    return None


def lk_roofline(tr):
    """Roofline entry of the matcher: algorithmic bytes per launch (one launch
    per frame step covers every scale) over its average measured duration."""
    k, prof = tr["kernels"].get("lk_track"), tr["prof"]
    if not k or not prof:
        return None
    avg_parts = [u / prof["frames"] for u in prof["per_scale"]]
    bytes_per_launch = lk_algorithmic_bytes(tr["shapes"], avg_parts)
    achieved = bytes_per_launch / (k["avg_us"] * 1e-6) / 1e9
    nscales = len(tr["shapes"])
    solves = sum(n * (nscales - s) for s, n in enumerate(avg_parts))
    flop_rate = solves * LK_FLOPS_PER_SOLVE / (k["avg_us"] * 1e-6) / 1e12
    fp32 = {"achieved_tflops": round(flop_rate, 3), "peak_tflops": FP32_PEAK_TFLOPS,
            "frac": round(flop_rate / FP32_PEAK_TFLOPS, 5), "level_solves_per_launch": round(solves, 1),
            "flops_per_level_solve": LK_FLOPS_PER_SOLVE,
            "model": "upper bound: every solve counted with 5 Gauss-Newton steps"}
    per_solve = measured_valu()
    valu_entry = None
    if per_solve:
        instr = per_solve * solves
        rate = instr / (k["avg_us"] * 1e-6)
        valu_entry = {"instr_per_level_solve": per_solve, "instr_per_launch_est": int(instr),
                      "achieved_ginstr_s": round(rate / 1e9, 1),
                      "peak_ginstr_s": round(VALU_ISSUE_PEAK / 1e9, 1), "frac": round(rate / VALU_ISSUE_PEAK, 4),
                      "source": "instructions per level solve from a separate rocprofv3 SQ_INSTS_VALU run of the same "
                                "workload (" + os.path.relpath(VALU_PROFILE, ROOT) + "), times this run's level "
                                "solves per launch, over this run's live launch time",
                      "model": "one wave64 VALU instruction per SIMD every 4 cycles, 256 CUs x 4 SIMDs at 2.4 GHz"}
    return {"kernel": "lk_track", "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(),
            "bytes_per_launch": int(bytes_per_launch), "avg_us": k["avg_us"], "fp32": fp32, "valu": valu_entry,
            "note": "the matcher is bound by VALU instruction issue, not HBM or FP32 throughput (see valu, fp32): "
                    "bytes = both FP32 frames of every scale once + 40 B per particle matched; traffic from "
                    "rocprofv3 FETCH_SIZE + WRITE_SIZE"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=270)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-micro", action="store_true")
    ap.add_argument("--no-sparse", action="store_true", help="skip the FAST th 20 run")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-stream batched step")
    ap.add_argument("--no-bc2s", action="store_true", help="skip the bc2s strategy run")
    ap.add_argument("--fast-th", type=float, default=FAST_TH)
    ap.add_argument("--prof-steps", type=int, default=30,
                    help="frame steps after the timed region run with per-kernel HIP events")
    ap.add_argument("--batch-tuning", type=str, default=None,
                    help="lk_ppw,blur_th for the 64-stream batch instead of the autotuner (counter runs)")
    ap.add_argument("--batch-no-sequential", action="store_true",
                    help="with --only batch: skip the sequential per-stream trackers (counter runs)")
    ap.add_argument("--only", choices=("micro", "batch"), default=None,
                    help="run only the Gaussian / pyramid microbenchmark or only the 64-stream batch and print "
                         "its JSON (profiling runs; not the benchmark line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    tuning = tuple(int(x) for x in args.batch_tuning.split(",")) if args.batch_tuning else None

    import cuimg_amd as cu
    ctx = cu.Context(0)
    if args.only == "micro":
        print(json.dumps({"gaussian_pyramid_roofline": gaussian_microbench(cu, ctx)}), flush=True)
        ctx.close()
        return
    if args.only == "batch":
        print(json.dumps({"batch_64_streams": batch_bench(cu, ctx, tuning, not args.batch_no_sequential)}),
              flush=True)
        ctx.close()
        return
    pct_steps = 50 if args.steps >= 50 else 0
    video, frames = upload_video(cu, ctx, args.warmup + args.steps + pct_steps + args.prof_steps)
    log(f"tracker benchmark, FAST th {args.fast_th:g}")
    tr = tracker_bench(cu, ctx, frames, args.fast_th, args.warmup, args.steps, args.prof_steps, pct_steps)
    log(f"  {tr['updates'] / tr['seconds']:.4g} particle-updates/s, {tr['frames'] / tr['seconds']:.1f} frames/s")
    sparse = None
    if not args.no_sparse:
        steps = min(args.steps, 120)
        log(f"tracker benchmark, FAST th {FAST_TH_SPARSE:g}")
        sp = tracker_bench(cu, ctx, frames, FAST_TH_SPARSE, min(args.warmup, 30), steps, 0)
        sparse = summary(sp)
        log(f"  {sparse['particle_updates_per_s']:.4g} particle-updates/s, {sparse['frames_per_s']} frames/s")
    for im in frames:
        im.free()
    bc2s, bc2s_video = None, None
    if not args.no_bc2s:
        # the integer descent follows at most 2 x 10 px per frame and the
        # reference's bc2s strategy has no coarse-to-fine prediction, so the
        # affine video (up to ~35 px at the frame corners) loses its finest
        # scale: the bc2s run uses the +-3 px subpixel translation generator
        steps = min(args.steps, 120)
        warm = min(args.warmup, 30)
        prof = min(args.prof_steps, 20)
        bc2s_video = cu.synth_video(SEED, NR, NC, warm + steps + prof, 0)
        bframes = [ctx.upload(bc2s_video[f]) for f in range(len(bc2s_video))]
        log(f"bc2s strategy, FAST th {args.fast_th:g}, detector every {BC2S_DET_FREQ} frames")
        bc = tracker_bench(cu, ctx, bframes, args.fast_th, warm, steps, prof, 0, profile="C", det_freq=BC2S_DET_FREQ,
                           kernel_names=BC2S_KERNELS)
        bc2s = summary(bc)
        bc2s["workload"] = (f"tracker<bc2s_fast_gradient_cpu> (the reference sample's strategy and settings): "
                            f"1920x1080 u8 video translated by a subpixel shift in [-3, 3] px per frame, {NSCALES} "
                            f"scales, FAST th {args.fast_th:g}, detector every {BC2S_DET_FREQ} frames, filtering "
                            f"every frame, {steps} steps")
        log(f"  {bc2s['particle_updates_per_s']:.4g} particle-updates/s, {bc2s['frames_per_s']} frames/s")
        for im in bframes:
            im.free()
    log("tracker from host frames")
    host_in = host_frames_bench(cu, ctx, video, args.fast_th, 20, min(args.steps, 100))
    log(f"  {host_in['frames_per_s']} frames/s from host memory")
    batch = None
    if not args.no_batch:
        log("batched 64-stream step")
        batch = batch_bench(cu, ctx, tuning)
        log(f"  batch {batch['batch_ms_per_step']} ms/step vs sequential {batch['sequential_ms_per_step']} ms/step "
            f"({batch['speedup']}x), identical {batch['streams_identical_to_own_tracker']}")
    micro = [] if args.no_micro else gaussian_microbench(cu, ctx)
    adjoint = None if args.no_micro else adjoint_bench(cu, ctx)
    if adjoint:
        log(f"adjoint: {adjoint['avg_us']} us per call ({adjoint['particles_per_s']:.4g} particles/s)")
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(video, args.cpu_budget, args.fast_th)
        if bc2s is not None:
            bc2s["cpu_baseline"] = bc2s_cpu_baseline(bc2s_video, min(args.cpu_budget, 8.0), args.fast_th)

    s = summary(tr)
    line = {
        "metric": METRIC,
        "value": s["particle_updates_per_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain pyramid-LK tracker, 1920x1080 u8 affine-motion video, 4 scales, "
                               f"sigma=1 7-tap FP32 blur, FAST th {args.fast_th:g} "
                               f"(~{s['updates_per_frame'] / 1000:.0f}k particle updates/frame), 7x7 LK <=5 it",
                   "frames": args.warmup + args.steps, "parallelism": "1 GPU"},
        "frames_per_s": s["frames_per_s"],
        "level_solves_per_s": s["level_solves_per_s"],
        "step_ms_events": s["step_ms_events"],
        "particles_per_scale": s["particles_per_scale"],
        "updates_per_frame": s["updates_per_frame"],
        "roofline": lk_roofline(tr),
        "kernels": s["kernels"],
        "fast_th_20": sparse,
        "host_frames": host_in,
        "batch_64_streams": batch,
        "bc2s": bc2s,
        "gaussian_pyramid_roofline": micro,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
