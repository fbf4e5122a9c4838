"""Pins the oracle's Gaussian coefficients and separable convolution.

Coefficients: against the reference's own generator output (committed fixture
and, where the reference tree is present, the freshly built generator) and the
closed form.  Convolution: impulse response, constant and ramp images, a
brute-force 2D clamp-to-edge sum, separability and an FP64 twin.
"""
import json
import math
import os
import re

import numpy as np
import pytest

from conftest import REFERENCE, ROOT

FIXTURE = os.path.join(ROOT, "tests", "golden", "meta_gaussian_coefs.json")


def test_meta_coefs_match_reference_generator_fixture(orc):
    data = json.load(open(FIXTURE))["coefs"]
    n = 0
    for sigma, table in data.items():
        for x, lit in table.items():
            want = np.float32(float(lit.rstrip("f")))
            got = np.float32(orc.coef_meta(int(sigma), int(x)))
            assert got == want, (sigma, x, got, want)
            n += 1
    assert n > 100


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not mounted")
def test_meta_coefs_match_freshly_built_generator(orc):
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    pat = re.compile(r"meta_kernel<meta_gaussian<0, (\d+)>, (-?\d+)>.*?return ([^;]+);")
    checked = 0
    for sigma in range(1, 51):
        path = os.path.join(ROOT, "oracle", "_ref", "meta_gaussian", f"meta_gaussian_coefs_{sigma}.h")
        for line in open(path):
            m = pat.search(line)
            if not m:
                continue
            x = int(m.group(2))
            want = np.float32(float(m.group(3).rstrip("f")))
            assert np.float32(orc.coef_meta(sigma, x)) == want, (sigma, x)
            checked += 1
    assert checked > 3000


@pytest.mark.parametrize("sigma", [1, 2, 4, 7])
def test_coefs_closed_form(orc, sigma):
    K = int(math.ceil(3 * sigma))
    for x in range(-K, K + 1):
        exact = math.exp(-x * x / (2.0 * sigma * sigma)) / (math.sqrt(2 * math.pi) * sigma)
        rt = orc.coef_runtime(sigma, x)
        assert abs(rt - exact) <= 4e-7 * exact + 1e-12
        meta = orc.coef_meta(sigma, x)
        # six significant digits
        assert abs(meta - exact) <= 6e-6 * exact
    c = orc.coefs(sigma, K)
    assert np.all(c == c[::-1])  # symmetric
    # not normalised: the truncated sum is below 1 (sigma=1, 7 taps ~ 0.9973)
    assert 0.99 < c.sum() < 1.0


def brute_2d(img, coef, K):
    """Direct 2D clamp-to-edge sum in float64: out = sum_ij c_i c_j in(r+i, c+j)."""
    nr, nc = img.shape
    out = np.zeros((nr, nc))
    for r in range(nr):
        for c in range(nc):
            s = 0.0
            for i in range(-K, K + 1):
                rr = min(max(r + i, 0), nr - 1)
                for j in range(-K, K + 1):
                    cc = min(max(c + j, 0), nc - 1)
                    s += float(coef[i + K]) * float(coef[j + K]) * float(img[rr, cc])
            out[r, c] = s
    return out


@pytest.mark.parametrize("sigma", [1, 2])
def test_impulse_response_is_outer_product(orc, sigma):
    K = 3 * sigma
    c = orc.coefs(sigma, K)
    img = np.zeros((4 * K + 3, 4 * K + 5), np.float32)
    r0, c0 = 2 * K + 1, 2 * K + 2
    img[r0, c0] = 1.0
    out = orc.gaussian(img, c, K)
    want = np.zeros_like(out)
    want[r0 - K:r0 + K + 1, c0 - K:c0 + K + 1] = np.outer(c, c).astype(np.float32)
    assert np.array_equal(out, want)


def test_constant_and_ramp(orc):
    K = 3
    c = orc.coefs(1, K)
    S = float(c.astype(np.float64).sum())
    img = np.full((20, 30), 100.0, np.float32)
    out = orc.gaussian(img, c, K)
    assert np.allclose(out, 100.0 * S * S, rtol=2e-6)
    ramp = np.tile(np.arange(30, dtype=np.float32) * 3.0, (20, 1))
    out = orc.gaussian(ramp, c, K)
    inner = out[:, K:-K]
    want = (ramp * S * S)[:, K:-K]
    assert np.allclose(inner, want, rtol=3e-6, atol=1e-4)


@pytest.mark.parametrize("shape,sigma", [((17, 13), 1), ((13, 17), 2), ((5, 4), 2), ((1, 9), 1), ((9, 1), 1)])
def test_brute_force_2d(orc, shape, sigma):
    rng = np.random.default_rng(7)
    K = 3 * sigma
    c = orc.coefs(sigma, K)
    img = rng.uniform(0, 255, shape).astype(np.float32)
    out = orc.gaussian(img, c, K)
    ref = brute_2d(img, c, K)
    assert np.allclose(out, ref, rtol=1e-6, atol=1e-4)


def test_separability_and_fp64_twin(orc):
    rng = np.random.default_rng(3)
    K = 6
    c = orc.coefs(2, K)
    img = rng.uniform(0, 255, (64, 80)).astype(np.float32)
    out = orc.gaussian(img, c, K)
    # transposing the image swaps the two passes
    out_t = orc.gaussian(np.ascontiguousarray(img.T), c, K).T
    assert np.allclose(out, out_t, rtol=1e-6)
    ref64 = orc.gaussian_f64(img.astype(np.float64), c.astype(np.float64), K)
    rel = np.abs(out - ref64) / np.maximum(np.abs(ref64), 1.0)
    assert rel.max() < 1e-6


def test_u8_input_equals_float_input(orc):
    rng = np.random.default_rng(11)
    img = rng.integers(0, 256, (31, 45), dtype=np.uint8)
    c = orc.coefs(1, 3)
    assert np.array_equal(orc.gaussian(img, c, 3), orc.gaussian(img.astype(np.float32), c, 3))


def test_rgba_channels_independent(orc):
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 255, (12, 10, 4)).astype(np.float32)
    c = orc.coefs(1, 3)
    out = orc.gaussian(img, c, 3)
    for k in range(4):
        assert np.array_equal(out[:, :, k], orc.gaussian(np.ascontiguousarray(img[:, :, k]), c, 3))


def test_fused_mode_matches_independent_restatement(orc):
    """The oracle's fused mode (the GPU's arithmetic: tap +K first, every
    further tap added with a fused multiply-add, row pass then column pass,
    clamped reads) against a Python restatement with exactly rounded FMAs."""
    from test_oracle_lk import f32_fma
    rng = np.random.default_rng(21)
    for shape, K, kind in (((9, 11), 3, "f32"), ((6, 5), 2, "u8"), ((4, 13), 4, "f32")):
        img = (rng.integers(0, 256, shape, dtype=np.uint8) if kind == "u8"
               else rng.uniform(0, 255, shape).astype(np.float32))
        coef = orc.coefs(1, K)
        nr, nc = shape
        x = img.astype(np.float32)
        tmp = np.zeros(shape, np.float32)
        for r in range(nr):
            for c in range(nc):
                s = np.float32(x[r, min(c + K, nc - 1)] * coef[2 * K])
                for j in range(K - 1, -K - 1, -1):
                    s = f32_fma(coef[j + K], x[r, min(max(c + j, 0), nc - 1)], s)
                tmp[r, c] = s
        want = np.zeros(shape, np.float32)
        for r in range(nr):
            for c in range(nc):
                s = np.float32(tmp[min(r + K, nr - 1), c] * coef[2 * K])
                for j in range(K - 1, -K - 1, -1):
                    s = f32_fma(coef[j + K], tmp[min(max(r + j, 0), nr - 1), c], s)
                want[r, c] = s
        got = orc.gaussian(img, coef, K, fused=True)
        assert np.array_equal(got, want), (shape, K, kind)
        # and it is a different rounding from the unfused reading
        assert np.abs(orc.gaussian(img, coef, K) - got).max() < 1e-3
