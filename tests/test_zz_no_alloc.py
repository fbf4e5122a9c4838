"""The allocation-free matcher entry point (cuimg_hip_lk_track_noalloc) against
cuimg_hip_lk_track, through the C ABI: the same particles after the move and
the same decisions, bit for bit, on two calls in a row (the second with fewer
particles, so stale scratch from the first would show), with decision buffers
passed in and with the fixed scratch holding them; an oversized call without
decision buffers is refused before any device work."""
import ctypes

import numpy as np
import pytest

import cuimg_amd as cu

pytestmark = pytest.mark.gpu

CAPACITY = 1 << 18


@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


def matcher_inputs(ctx):
    """The matcher images and one scale's particle vector of a tracker that
    ran two frames (particles of the first frame, images of the second)."""
    video = cu.synth_video(0xC0FFEE43, 240, 320, 2, 1)
    L = 3
    t = cu.Tracker(ctx, 240, 320, cu.profile_params("B", L, 8.0, with_merge=0, detector_frequency=100))
    t.run(video[0])
    before = [t.particles(s) for s in range(L)]
    t.run(video[1])
    imgs = [t.matcher_images(s) for s in range(L)]
    t.close()
    prev = [ctx.upload(a) for a, _ in imgs]
    cur = [ctx.upload(b) for _, b in imgs]
    return prev, cur, before


def call(ctx, fn, prev, cur, scale, parts, with_decisions, sparse_img=None):
    L = ctx.L
    n = len(parts)
    prm = cu.LkParams(7, 5, 0.1, 49, 300, 0)
    a = (cu.Img * len(prev))(*[x.d for x in prev])
    b = (cu.Img * len(cur))(*[x.d for x in cur])
    buf = cu.DeviceBuffer.from_array(ctx, parts)
    dec = [cu.DeviceBuffer(ctx, 4 * n), cu.DeviceBuffer(ctx, 8 * n), cu.DeviceBuffer(ctx, 4 * n)]
    ptrs = [d.p for d in dec] if with_decisions else [None, None, None]
    sp = ctypes.byref(sparse_img.d) if sparse_img is not None else None
    ctx.check(fn(ctx.h, a, b, len(prev), scale, buf.p, n, ctypes.byref(prm), sp, *ptrs), fn.__name__)
    out = buf.download(cu.PARTICLE_DTYPE, n)
    decisions = ((dec[0].download(np.int32, n), dec[1].download(np.float32, (n, 2)), dec[2].download(np.float32, n))
                 if with_decisions else None)
    for d in [buf] + dec:
        d.free()
    return out, decisions


@pytest.mark.parametrize("with_decisions", [True, False])
def test_noalloc_equals_allocating_entry_point(ctx, with_decisions):
    prev, cur, before = matcher_inputs(ctx)
    L = ctx.L
    for scale, take in ((0, None), (0, 300), (1, None)):
        parts = before[scale] if take is None else before[scale][:take]
        print(f"scale {scale}, {len(parts)} particles, decisions {'given' if with_decisions else 'in scratch'}",
              flush=True)
        assert len(parts) > 100
        sp_a = ctx.image(*prev[scale].shape, cu.U32)
        sp_b = ctx.image(*prev[scale].shape, cu.U32)
        empty = np.full(prev[scale].shape, 0xFFFFFFFF, np.uint32)
        sp_a.upload(empty)
        sp_b.upload(empty)
        want, dw = call(ctx, L.cuimg_hip_lk_track, prev, cur, scale, parts, with_decisions, sp_a)
        got, dg = call(ctx, L.cuimg_hip_lk_track_noalloc, prev, cur, scale, parts, with_decisions, sp_b)
        assert got.tobytes() == want.tobytes()
        assert np.array_equal(sp_a.download(), sp_b.download())
        if with_decisions:
            assert all(np.array_equal(x, y) for x, y in zip(dw, dg))
        assert (got["age"] > 0).sum() > 50


def test_noalloc_rejects_more_particles_than_its_scratch(ctx):
    prev, cur, _ = matcher_inputs(ctx)
    L = ctx.L
    prm = cu.LkParams(7, 5, 0.1, 49, 300, 0)
    a = (cu.Img * 3)(*[x.d for x in prev])
    b = (cu.Img * 3)(*[x.d for x in cur])
    buf = cu.DeviceBuffer(ctx, 64)
    r = L.cuimg_hip_lk_track_noalloc(ctx.h, a, b, 3, 0, buf.p, CAPACITY + 1, ctypes.byref(prm), None, None, None,
                                     None)
    assert r == -4  # CUIMG_E_CAPACITY, refused before any launch
    buf.free()
