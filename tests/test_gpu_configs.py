"""GPU parity at the benchmark workloads' exact shapes.

  * Gaussian roofline workload: float32 4096x4096 and float32x4 3840x2160
    images, i.i.d. U[0,255) from a seed, sigma 1, 2, 4 with 2*3*sigma+1 taps,
    clamp-to-edge: within 1e-5 relative of the oracle and bit-exact against its
    fused mode (the GPU's arithmetic); and the 5-level 2x box
    pyramid of the 4096x4096 image: bit-exact.  The blur's tile height is
    picked from the image size, so these shapes run tile shapes the smaller
    cases do not.
  * Batched workload: 64 streams of width U[320,1920], height round(9w/16), 3
    or 4 scales, U[1k,150k] seeded particles, 8 frames of the affine-motion
    video: every stream of the batch equals its own tracker, bit for bit, on
    every frame (particles, decisions) and at the end (sparse image,
    saliency, matches).  Own trackers are checked against the oracle by the
    lockstep tests; together this is each stream's parity with the oracle.
"""
import os
import sys

import numpy as np
import pytest

import cuimg_amd as cu

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


def rel_err(got, ref):
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max() + 1e-30)
    return float((np.abs(got.astype(np.float64) - ref) / scale).max())


@pytest.fixture(scope="module")
def images():
    rng = np.random.default_rng(3)
    return {"f32": rng.uniform(0, 255, (4096, 4096)).astype(np.float32),
            "f32x4": rng.uniform(0, 255, (2160, 3840, 4)).astype(np.float32)}


@pytest.mark.parametrize("sigma", [1, 2, 4])
@pytest.mark.parametrize("kind", ["f32", "f32x4"])
def test_gaussian_roofline_shapes(ctx, orc, images, kind, sigma):
    img = images[kind]
    print(f"gaussian {kind} {img.shape} sigma={sigma}", flush=True)
    K = 3 * sigma
    coef = orc.coefs(sigma, K)
    src = ctx.upload(img)
    dst = ctx.image(img.shape[0], img.shape[1], cu.F32 if kind == "f32" else cu.F32X4)
    ctx.gaussian(src, dst, sigma, K)
    got = dst.download()
    ref = orc.gaussian(img, coef, K)
    err = rel_err(got, ref)
    print(f"  max relative error {err:.3g}", flush=True)
    assert err < 1e-5
    if K < 9:  # the vector kernel: the oracle's fused mode bit for bit (19+ taps run on the matrix cores)
        assert np.array_equal(got, orc.gaussian(img, coef, K, fused=True))
    src.free()
    dst.free()


def test_pyramid_roofline_shape(ctx, orc, images):
    """The 5-level pyramid of the microbenchmark, built in place on level 0."""
    img = images["f32"]
    levels = [ctx.upload(img)]
    ref = [img]
    for _ in range(1, 5):
        r, c = levels[-1].shape
        levels.append(ctx.image(r // 2, c // 2, cu.F32))
        ref.append(orc.pyr_down(ref[-1]))
    ctx.pyramid(levels[0], levels)
    for l in range(5):
        assert np.array_equal(levels[l].download(), ref[l]), l
    for im in levels:
        im.free()


@pytest.mark.parametrize("fast_th", [256.0, 20.0])
def test_batched_workload_streams_equal_own_trackers(ctx, fast_th):
    """fast_th 256 is the benchmark's (no saliency reaches it: the counts are
    the placement's); 20 adds detected particles every frame on top."""
    shapes, params, parts = bench.batch_workload(cu)
    n = len(shapes)
    assert n == 64
    for p in params:
        p.fast_th = fast_th
    print(f"batched workload: {n} streams, {sum(h * w for h, w in shapes) / 1e6:.1f} Mpx, "
          f"{sum(len(p[0]) for p in parts)} particles at scale 0, fast_th {fast_th}", flush=True)
    nframes = bench.BATCH_FRAMES
    frames = [[None] * n for _ in range(nframes)]
    for i, (h, w) in enumerate(shapes):
        v = cu.synth_video(bench.BATCH_SEED + 1 + i, h, w, nframes, 1)
        for f in range(nframes):
            frames[f][i] = ctx.upload(v[f])
    batch = cu.Batch(ctx, shapes, params)
    solo = [cu.Tracker(ctx, h, w, params[i]) for i, (h, w) in enumerate(shapes)]
    views = [batch.tracker(i) for i in range(n)]
    live = 0
    for f in range(nframes):
        print(f"frame {f}", flush=True)
        batch.step(frames[f])
        for i in range(n):
            solo[i].run(frames[f][i])
        if f == 0:
            for i in range(n):
                for s, p in enumerate(parts[i]):
                    views[i].set_particles(s, p)
                    solo[i].set_particles(s, p)
            continue
        for i in range(n):
            for s in range(params[i].nscales):
                a, b = views[i].particles(s), solo[i].particles(s)
                assert a.tobytes() == b.tobytes(), (f, i, s)
                da, db = views[i].decisions(s), solo[i].decisions(s)
                assert all(np.array_equal(x, y) for x, y in zip(da, db)), (f, i, s)
                live += int((da[0] == 1).sum())
    for i in range(n):
        for s in range(params[i].nscales):
            assert np.array_equal(views[i].sparse(s), solo[i].sparse(s)), (i, s)
            assert np.array_equal(views[i].saliency(s), solo[i].saliency(s)), (i, s)
            assert views[i].compact_has_run(s) == solo[i].compact_has_run(s)
            if solo[i].compact_has_run(s):
                assert np.array_equal(views[i].matches(s), solo[i].matches(s)), (i, s)
    print(f"  {live} accepted particle updates compared", flush=True)
    assert live > 1_000_000
    for t in solo:
        t.close()
    batch.close()
    for row in frames:
        for im in row:
            im.free()
