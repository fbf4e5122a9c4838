"""Pins the oracle's box pyramid and FAST saliency.

Pyramid: integer closed form (floor of the 2x2 mean), odd sizes, float order.
FAST: known-answer arcs (lengths 8/9/12/16, wrap-around through index 0),
threshold equality, the modulo-256 score, and agreement with an independent
restatement (circular runs over a doubled ring) on random and structured
patches.
"""
import numpy as np
import pytest

RING = [(-3, 0), (-3, 1), (-2, 2), (-1, 3), (0, 3), (1, 3), (2, 2), (3, 1),
        (3, 0), (3, -1), (2, -2), (1, -3), (0, -3), (-1, -3), (-2, -2), (-3, -1)]


@pytest.mark.parametrize("shape", [(8, 8), (9, 7), (240, 320), (135, 241)])
def test_pyr_down_u8_closed_form(orc, shape):
    rng = np.random.default_rng(shape[0])
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    out = orc.pyr_down(img)
    h, w = shape[0] // 2, shape[1] // 2
    a = img[:2 * h, :2 * w].astype(np.int32)
    want = (a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2]) // 4
    assert out.shape == (h, w)
    assert np.array_equal(out, want.astype(np.uint8))


def test_pyr_down_f32_order(orc):
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 255, (33, 20)).astype(np.float32)
    out = orc.pyr_down(img)
    a = img[:32, :20]
    s = (a[0::2, 0::2] + a[0::2, 1::2]).astype(np.float32)
    s = (s + a[1::2, 0::2]).astype(np.float32)
    s = (s + a[1::2, 1::2]).astype(np.float32)
    assert np.array_equal(out, (s / np.float32(4)).astype(np.float32))


def py_fast(img, r, c, n, th):
    """Independent restatement: maximal circular run of ring points beyond th
    with one sign; score = max(bright sum, dark sum) mod 256."""
    nr, nc = img.shape
    cl = lambda y, x: int(img[min(max(y, 0), nr - 1), min(max(x, 0), nc - 1)])
    vp = cl(r, c)
    states, bright, dark = [], 0, 0
    for dr, dc in RING:
        d = cl(r + dr, c + dc) - vp
        if abs(d) > th:
            states.append(1 if d > 0 else -1)
            if d > 0:
                bright += abs(d)
            else:
                dark += abs(d)
        else:
            states.append(0)
    if states[0] != 0 and all(s == states[0] for s in states):
        run = 16
    else:
        run, cur, prev = 0, 0, 0
        for s in states + states:
            cur = cur + 1 if (s != 0 and s == prev) else (1 if s != 0 else 0)
            prev = s
            run = max(run, cur)
        run = min(run, 16)
    if run < n:
        return 0
    return max(bright, dark) % 256


def arc_patch(start, length, delta, center=100, noise=None):
    img = np.full((7, 7), center, np.uint8)
    for k in range(length):
        dr, dc = RING[(start + k) % 16]
        img[3 + dr, 3 + dc] = center + delta
    return img


@pytest.mark.parametrize("length,score", [(8, 0), (9, 9 * 50 % 256), (12, 12 * 50 % 256), (16, 16 * 50 % 256)])
@pytest.mark.parametrize("start", [0, 5, 12])
def test_known_answer_arcs(orc, length, score, start):
    assert orc.fast_score(arc_patch(start, length, 50), 3, 3, 9, 20) == score
    assert orc.fast_score(arc_patch(start, length, -50), 3, 3, 9, 20) == score


def test_threshold_equality_and_wrap(orc):
    assert orc.fast_score(arc_patch(0, 10, 20), 3, 3, 9, 20) == 0      # dist == th is not beyond
    assert orc.fast_score(arc_patch(0, 10, 21), 3, 3, 9, 20) == 210
    # wrap-around: arc 13,14,15,0..5 (9 points) crossing index 0
    assert orc.fast_score(arc_patch(13, 9, 30), 3, 3, 9, 20) == 270 % 256
    # the score wraps modulo 256: 16 * 155 = 2480 -> 176
    assert orc.fast_score(arc_patch(0, 16, 155), 3, 3, 9, 20) == 2480 % 256
    # mixed signs: bright 9-arc, dark rest; score = max of sums
    img = arc_patch(0, 9, 60)
    for k in range(9, 16):
        dr, dc = RING[k]
        img[3 + dr, 3 + dc] = 30
    assert orc.fast_score(img, 3, 3, 9, 20) == max(9 * 60, 7 * 70) % 256


@pytest.mark.parametrize("n", [9, 12])
def test_random_patches_agree_with_independent_restatement(orc, n):
    rng = np.random.default_rng(100 + n)
    mism = 0
    for trial in range(3000):
        kind = trial % 3
        if kind == 0:
            img = rng.integers(0, 256, (7, 7), dtype=np.uint8)
        elif kind == 1:  # structured: random arc of random length over noise
            img = arc_patch(int(rng.integers(16)), int(rng.integers(6, 17)),
                            int(rng.choice([-1, 1]) * rng.integers(15, 120)), center=int(rng.integers(120, 136)))
            img = np.clip(img.astype(int) + rng.integers(-6, 7, (7, 7)), 0, 255).astype(np.uint8)
        else:  # two-level image (step edges / corners)
            img = np.where(rng.random((7, 7)) < 0.5, rng.integers(0, 256), rng.integers(0, 256)).astype(np.uint8)
        th = int(rng.integers(0, 60))
        if orc.fast_score(img, 3, 3, n, th) != py_fast(img, 3, 3, n, th):
            mism += 1
    assert mism == 0


def test_saliency_respects_mask_and_border(orc):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (40, 50), dtype=np.uint8)
    mask = np.full(img.shape, 255, np.uint8)
    mask[10:20, 10:20] = 0
    sal = orc.fast_saliency(img, mask, 9, 20)
    assert np.all(sal[10:20, 10:20] == 0)
    for r in range(40):
        for c in range(50):
            want = 0 if mask[r, c] == 0 else py_fast(img, r, c, 9, 20)
            assert sal[r, c] == want


def test_pyr_down_gauss_sampling(orc):
    """Blur-then-subsample level: a constant image gives c * sum(coef)^2, and an
    impulse at (r0, c0) gives coef[2i - r0 + 6] * coef[2j - c0 + 6] (the
    separable impulse response sampled at even pixels)."""
    cf = orc.coefs(1, 6).astype(np.float64)
    out = orc.pyr_down_gauss(np.full((40, 36), 10.0, np.float32))
    assert out.shape == (20, 18)
    assert np.allclose(out, 10.0 * cf.sum() ** 2, rtol=1e-6)
    for r0, c0 in [(20, 18), (21, 17)]:
        img = np.zeros((41, 37), np.float32)
        img[r0, c0] = 1.0
        out = orc.pyr_down_gauss(img).astype(np.float64)
        want = np.zeros_like(out)
        for i in range(out.shape[0]):
            for j in range(out.shape[1]):
                a, b = 2 * i - r0 + 6, 2 * j - c0 + 6
                if 0 <= a <= 12 and 0 <= b <= 12:
                    want[i, j] = cf[a] * cf[b]
        assert np.allclose(out, want, rtol=1e-6, atol=1e-12)
