"""GPU parity of the image kernels against the CPU oracle (through the C ABI).

Gaussian: within 1e-5 relative of the oracle's reading of the reference (each
product rounded, taps accumulated from +K down).  The vector kernel (u8 input,
or FP32 / FP32x4 with fewer than 19 taps) is also bit-exact against the
oracle's fused mode (the same taps added with fused multiply-adds, the GPU's
arithmetic); the matrix-core kernel (FP32 / FP32x4, 19 taps and more) adds each
output's products in the matrix unit's order, so only the relative bar applies
to it.
Pyramid, border fill and FAST saliency: bit-exact.
"""
import numpy as np
import pytest

import cuimg_amd as cu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


def rel_err(got, ref):
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max() + 1e-30)
    return float((np.abs(got.astype(np.float64) - ref) / scale).max())


MFMA_MIN_K = 9  # FP32 / FP32x4 half widths from which the matrix-core kernel runs

GAUSS_CASES = [
    ("u8", (240, 320), 1),
    ("u8", (1080, 1920), 1),
    ("u8", (135, 241), 2),
    ("f32", (17, 13), 1),
    ("f32", (5, 4), 2),
    ("f32", (1, 9), 1),
    ("f32", (9, 1), 1),
    ("f32", (1024, 1024), 1),
    ("f32", (1024, 1000), 2),
    ("f32", (777, 1031), 4),
    ("f32x4", (270, 480), 1),
    ("f32x4", (135, 250), 4),
    # full-width strips (the direct-to-LDS row pipeline), short images
    ("f32", (2048, 512), 4),
    ("f32", (5, 512), 1),
    ("f32", (3, 256), 3),
    ("f32", (40, 1024), 5),
    ("f32x4", (200, 640), 2),
    ("f32x4", (64, 128), 4),
    # outputs above 32 MiB: non-temporal stores
    ("f32", (2500, 3584), 1),
    ("f32x4", (1100, 2048), 2),
    ("u8", (3000, 3000), 1),
]


@pytest.mark.parametrize("kind,shape,sigma", GAUSS_CASES)
def test_gaussian_matches_oracle(ctx, orc, kind, shape, sigma):
    print(f"gaussian {kind} {shape} sigma={sigma}", flush=True)
    rng = np.random.default_rng(hash((kind, shape, sigma)) & 0xFFFF)
    K = 3 * sigma
    coef = orc.coefs(sigma, K)
    assert np.array_equal(coef, cu.gaussian_coefs(sigma, K))
    if kind == "u8":
        img = rng.integers(0, 256, shape, dtype=np.uint8)
        out_dt = cu.F32
    elif kind == "f32":
        img = rng.uniform(0, 255, shape).astype(np.float32)
        out_dt = cu.F32
    else:
        img = rng.uniform(0, 255, shape + (4,)).astype(np.float32)
        out_dt = cu.F32X4
    src = ctx.upload(img, border=3 if kind == "u8" else 0)
    dst = ctx.image(shape[0], shape[1], out_dt)
    ctx.gaussian(src, dst, sigma, K)
    got = dst.download()
    ref = orc.gaussian(img, coef, K)
    assert rel_err(got, ref) < 1e-5
    if kind == "u8" or K < MFMA_MIN_K:
        # the same operations as the oracle's fused mode: bit-exact
        assert np.array_equal(got, orc.gaussian(img, coef, K, fused=True))


def test_gaussian_multi_levels(ctx, orc):
    rng = np.random.default_rng(5)
    shapes = [(1080, 1920), (540, 960), (270, 480), (135, 240)]
    imgs = [rng.integers(0, 256, s, dtype=np.uint8) for s in shapes]
    srcs = [ctx.upload(i, border=3) for i in imgs]
    dsts = [ctx.image(s[0], s[1], cu.F32) for s in shapes]
    ctx.gaussian_multi(srcs, dsts, 1, 3)
    coef = orc.coefs(1, 3)
    for i, d in zip(imgs, dsts):
        assert rel_err(d.download(), orc.gaussian(i, coef, 3)) < 1e-5


@pytest.mark.parametrize("th", [1, 2, 16, 64, 128])
@pytest.mark.parametrize("width", [700, 768])
def test_gaussian_tile_heights_agree(ctx, orc, th, width):
    rng = np.random.default_rng(th)
    img = rng.uniform(0, 255, (300, width)).astype(np.float32)
    coef = orc.coefs(2, 6)
    src = ctx.upload(img)
    dst = ctx.image(300, width, cu.F32)
    ctx.convolve(src, dst, coef, 6, th=th)
    assert rel_err(dst.download(), orc.gaussian(img, coef, 6)) < 1e-5


def test_gaussian_tile_directions_bit_identical(ctx):
    """Odd tile rows stream bottom-up; every output must carry the same bits as
    when a single top-down tile computes it."""
    rng = np.random.default_rng(11)
    img = rng.uniform(0, 255, (300, 768)).astype(np.float32)
    src = ctx.upload(img)
    one, alt = ctx.image(300, 768, cu.F32), ctx.image(300, 768, cu.F32)
    for sigma in (1, 3):
        coef = cu.gaussian_coefs(sigma, 3 * sigma)
        ctx.convolve(src, one, coef, 3 * sigma, th=300)
        ctx.convolve(src, alt, coef, 3 * sigma, th=16)
        assert np.array_equal(one.download(), alt.download()), sigma


@pytest.mark.parametrize("kind", ["f32", "f32x4"])
def test_gaussian_matrix_core_tiles_agree(ctx, orc, kind):
    """The matrix-core kernel (25 taps): every output comes from the same
    16-row and 16-column blocks whatever the tile height, so all tile heights
    give the same bits; a two-image launch matches too."""
    rng = np.random.default_rng(21)
    shape = (301, 777) if kind == "f32" else (133, 301)
    img = rng.uniform(-50, 255, shape + ((4,) if kind == "f32x4" else ())).astype(np.float32)
    dt = cu.F32 if kind == "f32" else cu.F32X4
    coef = orc.coefs(4, 12)
    src = ctx.upload(img)
    outs = []
    for th in (1, 16, 48, 300):
        dst = ctx.image(shape[0], shape[1], dt)
        ctx.convolve(src, dst, coef, 12, th=th)
        outs.append(dst.download())
        dst.free()
    assert rel_err(outs[0], orc.gaussian(img, coef, 12)) < 1e-5
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    img2 = rng.uniform(0, 255, (97, 130) + ((4,) if kind == "f32x4" else ())).astype(np.float32)
    src2 = ctx.upload(img2)
    d1, d2 = ctx.image(shape[0], shape[1], dt), ctx.image(97, 130, dt)
    ctx.gaussian_multi([src, src2], [d1, d2], 4, 12)
    assert np.array_equal(d1.download(), outs[0])
    assert rel_err(d2.download(), orc.gaussian(img2, coef, 12)) < 1e-5
    for im in (src, src2, d1, d2):
        im.free()


def clamp_pad(a, b):
    return np.pad(a, b, mode="edge")


@pytest.mark.parametrize("shape,nlev", [((1080, 1920), 4), ((240, 320), 3), ((135, 241), 3), ((67, 131), 5),
                                        ((640, 960), 6), ((1000, 1300), 7), ((6000, 6000), 3)])
def test_pyramid_u8_bit_exact(ctx, orc, shape, nlev):
    rng = np.random.default_rng(shape[1])
    frame = rng.integers(0, 256, shape, dtype=np.uint8)
    f = ctx.upload(frame)
    levels, ref = [], [frame]
    r, c = shape
    for l in range(nlev):
        levels.append(ctx.image(r, c, cu.U8, border=3))
        r, c = r // 2, c // 2
    for l in range(1, nlev):
        ref.append(orc.pyr_down(ref[-1]))
    ctx.pyramid(f, levels)
    for l in range(nlev):
        got = levels[l].download_with_border()
        assert np.array_equal(got, clamp_pad(ref[l], 3)), l


def test_pyramid_f32_in_place(ctx, orc):
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 255, (512, 768)).astype(np.float32)
    levels = [ctx.upload(img)]
    ref = [img]
    for l in range(1, 5):
        r, c = levels[-1].shape
        levels.append(ctx.image(r // 2, c // 2, cu.F32))
        ref.append(orc.pyr_down(ref[-1]))
    ctx.pyramid(levels[0], levels)
    for l in range(5):
        assert np.array_equal(levels[l].download(), ref[l]), l


@pytest.mark.parametrize("shape,nlev,in_place", [((3072, 3072), 4, False), ((1024, 1536), 7, True), ((96, 4000), 5, False),
                                                 ((300, 1000), 6, True)])
def test_pyramid_f32_tilings(ctx, orc, shape, nlev, in_place):
    """Tile shapes from 64x64 to 256x16, ragged edges, and (3072^2) the
    non-temporal stores of images above 32 MiB."""
    print(f"pyramid f32 {shape} {nlev} levels in_place={in_place}", flush=True)
    rng = np.random.default_rng(shape[0] + nlev)
    img = rng.uniform(0, 255, shape).astype(np.float32)
    frame = ctx.upload(img)
    levels = [frame if in_place else ctx.image(shape[0], shape[1], cu.F32)]
    ref = [img]
    for l in range(1, nlev):
        r, c = levels[-1].shape
        levels.append(ctx.image(r // 2, c // 2, cu.F32))
        ref.append(orc.pyr_down(ref[-1]))
    ctx.pyramid(frame, levels)
    for l in range(nlev):
        assert np.array_equal(levels[l].download(), ref[l]), l


def test_pyr_down_leaves_borders(ctx, orc):
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (50, 70), dtype=np.uint8)
    src = ctx.upload(img, border=3)
    dst = ctx.image(25, 35, cu.U8, border=3)
    ctx.pyr_down(src, dst)
    assert np.array_equal(dst.download(), orc.pyr_down(img))


@pytest.mark.parametrize("shape", [(2, 2), (3, 5), (17, 13), (50, 70), (131, 257), (1080, 1920), (4096, 4096)])
@pytest.mark.parametrize("kind", ["u8", "f32", "f32_unaligned"])
def test_pyr_down_gauss_decimate(ctx, orc, shape, kind):
    """F32 images stream strips of 256 input columns (16-byte row loads when
    aligned, clamped 4-byte loads for the last partial strip and unaligned
    images); U8 images take the block kernel."""
    print(f"gauss decimate {kind} {shape}", flush=True)
    rng = np.random.default_rng(shape[0] * 31 + shape[1])
    if kind == "u8":
        img = rng.integers(0, 256, shape, dtype=np.uint8)
        src = ctx.upload(img, border=3)
    else:
        img = rng.uniform(0, 255, shape).astype(np.float32)
        # a one-pixel border puts the first pixel 4 bytes past a 16-byte boundary
        src = ctx.upload(img, border=1 if kind == "f32_unaligned" else 0)
    dst = ctx.image(shape[0] // 2, shape[1] // 2, cu.F32)
    ctx.pyr_down(src, dst, cu.PYR_GAUSS_DECIMATE)
    got = dst.download()
    assert rel_err(got, orc.pyr_down_gauss(img)) < 1e-5
    # the taps in the oracle's fused order: bit-exact
    g = orc.gaussian(img, orc.coefs(1, 6), 6, fused=True)
    assert np.array_equal(got, g[0:2 * (shape[0] // 2):2, 0:2 * (shape[1] // 2):2])


def test_pyr_down_mode_errors(ctx):
    src = ctx.upload(np.zeros((20, 20), np.uint8))
    with pytest.raises(cu.CuimgError):  # U8 output
        ctx.pyr_down(src, ctx.image(10, 10, cu.U8), cu.PYR_GAUSS_DECIMATE)
    with pytest.raises(cu.CuimgError):  # wrong size
        ctx.pyr_down(src, ctx.image(9, 10, cu.F32), cu.PYR_GAUSS_DECIMATE)
    with pytest.raises(cu.CuimgError):
        ctx.pyr_down(src, ctx.image(10, 10, cu.U8), 7)


@pytest.mark.parametrize("dtype", [cu.U8, cu.U32, cu.F32X2, cu.F32X4])
def test_fill_border_clamp(ctx, dtype):
    np_t, tail = {cu.U8: (np.uint8, ()), cu.U32: (np.uint32, ()), cu.F32X2: (np.float32, (2,)),
                  cu.F32X4: (np.float32, (4,))}[dtype]
    rng = np.random.default_rng(dtype)
    a = (rng.integers(0, 200, (13, 17) + tail)).astype(np_t)
    im = ctx.image(13, 17, dtype, border=4)
    im.upload(a)
    ctx.fill_border(im)
    pad = [(4, 4), (4, 4)] + [(0, 0)] * len(tail)
    assert np.array_equal(im.download_with_border(), np.pad(a, pad, mode="edge"))


@pytest.mark.parametrize("seed,shape", [(1, (240, 320)), (2, (1080, 1920)), (3, (61, 95))])
def test_fast_saliency_bit_exact(ctx, orc, seed, shape):
    frames = cu.synth_video(0xC0FFEE00 + seed, shape[0], shape[1], 1, 1)
    img = frames[0]
    rng = np.random.default_rng(seed)
    mask = np.where(rng.random(shape) < 0.2, 0, 255).astype(np.uint8)
    src = ctx.upload(img, border=3)
    ctx.fill_border(src)
    m = ctx.upload(mask)
    out = ctx.image(shape[0], shape[1], cu.U8)
    for th in (20, 5):
        ctx.fast(src, m, out, 9, th)
        assert np.array_equal(out.download(), orc.fast_saliency(img, mask, 9, th))
        ctx.fast(src, None, out, 9, th)
        assert np.array_equal(out.download(), orc.fast_saliency(img, None, 9, th))


def test_fast_small_border_uses_clamp(ctx, orc):
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (40, 37), dtype=np.uint8)
    src = ctx.upload(img, border=0)
    out = ctx.image(40, 37, cu.U8)
    ctx.fast(src, None, out, 9, 10)
    assert np.array_equal(out.download(), orc.fast_saliency(img, None, 9, 10))
