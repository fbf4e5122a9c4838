"""Adjoint of the blurred photometric LK loss.

The FP64 oracle's analytic gradient is pinned by central finite differences
(1e-3 px on d_p, 1e-2 gray levels on I1).  The GPU's FP32 adjoint must agree
with the FP64 oracle within 1e-4 relative (normwise over all particles / all
pixels) on the 1280x720, 50k-particle workload, and per particle away from
cancellation.
"""
import numpy as np
import pytest

import cuimg_amd as cu

GRID = 2.0 ** -12  # positions and displacements on this grid are exact in FP32


def workload(seed, nr, nc, n, margin=8, synth=True):
    rng = np.random.default_rng(seed)
    if synth:
        v = cu.synth_video(0xC0FFEE05 + seed, nr, nc, 2, 1)
        I0, I1 = v[0].astype(np.float32), v[1].astype(np.float32)
    else:
        I0 = rng.uniform(0, 255, (nr, nc)).astype(np.float32)
        I1 = rng.uniform(0, 255, (nr, nc)).astype(np.float32)
    pos = np.stack([rng.uniform(margin, nr - 1 - margin, n), rng.uniform(margin, nc - 1 - margin, n)], 1)
    pos = (np.round(pos / GRID) * GRID).astype(np.float32)
    d = (np.round(rng.uniform(-2, 2, (n, 2)) / GRID) * GRID).astype(np.float32)
    return I0, I1, pos, d


def keep_away_from_cell_edges(pos, d, eps=0.01):
    f1 = np.abs(((pos + d) % 1.0) - 0.5)
    f0 = np.abs((pos % 1.0) - 0.5)
    return (f1 < 0.5 - eps).all(1) & (f0 < 0.5 - eps).all(1)


def test_oracle_gradient_matches_finite_differences(orc):
    I0, I1, pos, d = workload(1, 48, 64, 25, margin=6)
    coef = orc.coefs(1, 3)
    ok = keep_away_from_cell_edges(pos, d)
    pos, d = pos[ok], d[ok]
    L, gd, gI = orc.loss_grad(I0, I1, coef, 3, pos, d)
    I0d, I1d = I0.astype(np.float64), I1.astype(np.float64)
    posd, dd = pos.astype(np.float64), d.astype(np.float64)
    assert abs(orc.loss_f64(I0d, I1d, coef, 3, posd, dd) - L) <= 1e-9 * abs(L)
    h = 1e-3
    for p in range(len(pos)):
        for k in range(2):
            dp, dm = dd.copy(), dd.copy()
            dp[p, k] += h
            dm[p, k] -= h
            fd = (orc.loss_f64(I0d, I1d, coef, 3, posd, dp) - orc.loss_f64(I0d, I1d, coef, 3, posd, dm)) / (2 * h)
            assert abs(fd - gd[p, k]) <= 1e-5 * max(1.0, abs(gd[p, k])), (p, k, fd, gd[p, k])
    rng = np.random.default_rng(7)
    hI = 1e-2
    pix = [(0, 0), (47, 63), (0, 30)] + [tuple(x) for x in rng.integers(0, (48, 64), (40, 2))]
    for (r, c) in pix:
        Ip, Im = I1d.copy(), I1d.copy()
        Ip[r, c] += hI
        Im[r, c] -= hI
        fd = (orc.loss_f64(I0d, Ip, coef, 3, posd, dd) - orc.loss_f64(I0d, Im, coef, 3, posd, dd)) / (2 * hI)
        assert abs(fd - gI[r, c]) <= 1e-6 * max(1.0, abs(gI[r, c])) + 1e-7 * np.abs(gI).max(), (r, c, fd, gI[r, c])


def test_oracle_zero_motion_zero_loss(orc):
    I0, _, pos, d = workload(2, 40, 50, 10)
    L, gd, gI = orc.loss_grad(I0, I0, orc.coefs(1, 3), 3, pos, np.zeros_like(d))
    assert L == 0.0 and not gd.any() and not gI.any()


def rel_norm(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.gpu
def test_gpu_adjoint_matches_fp64(orc):
    print("adjoint 1280x720, 50k particles", flush=True)
    I0, I1, pos, d = workload(5, 720, 1280, 50000)
    coef = orc.coefs(1, 3)
    L64, gd64, gI64 = orc.loss_grad(I0, I1, coef, 3, pos, d)
    ctx = cu.Context(0)
    a, b = ctx.upload(I0), ctx.upload(I1)
    L32, lp, gd, gI = cu.lk_loss_adjoint(ctx, a, b, pos, d, 1, 3, 7)
    print(f"L rel {abs(L32 - L64) / L64:.2e}  dL/dd rel {rel_norm(gd, gd64):.2e}  dL/dI1 rel {rel_norm(gI, gI64):.2e}")
    assert abs(L32 - L64) <= 1e-4 * L64
    assert rel_norm(gd, gd64) <= 1e-4
    assert rel_norm(gI, gI64) <= 1e-4
    # per-particle losses against the FP64 forward
    _, _, _ = orc.loss_grad(I0, I1, coef, 3, pos[:1], d[:1], want_dI=False)
    lp64 = np.array([orc.loss_grad(I0, I1, coef, 3, pos[i:i + 1], d[i:i + 1], want_dI=False)[0]
                     for i in range(0, 50000, 997)])
    assert np.abs(lp[::997] - lp64).max() <= 1e-4 * lp64.max()
    # per particle, away from cancellation
    mag = np.abs(gd64).max(1)
    big = mag > 1e-3 * mag.max()
    per = np.abs(gd - gd64).max(1)[big] / mag[big]
    assert (per <= 1e-4).mean() >= 0.999
    ctx.close()


@pytest.mark.gpu
def test_gpu_adjoint_edges_and_clamp(orc):
    """Particles whose windows reach and cross the image edge (the clamped path)."""
    rng = np.random.default_rng(9)
    nr, nc = 40, 56
    I0 = rng.uniform(0, 255, (nr, nc)).astype(np.float32)
    I1 = rng.uniform(0, 255, (nr, nc)).astype(np.float32)
    n = 300
    pos = np.stack([rng.uniform(0, nr - 1, n), rng.uniform(0, nc - 1, n)], 1)
    pos = (np.round(pos / GRID) * GRID).astype(np.float32)
    d = (np.round(rng.uniform(-2, 2, (n, 2)) / GRID) * GRID).astype(np.float32)
    d = np.where(pos + d < 0, -pos * 0.5, d).astype(np.float32)  # keep samples' anchor >= 0 half the time
    coef = orc.coefs(1, 3)
    L64, gd64, gI64 = orc.loss_grad(I0, I1, coef, 3, pos, d)
    ctx = cu.Context(0)
    L32, lp, gd, gI = cu.lk_loss_adjoint(ctx, ctx.upload(I0), ctx.upload(I1), pos, d, 1, 3, 7)
    assert abs(L32 - L64) <= 1e-4 * L64
    assert rel_norm(gd, gd64) <= 1e-4
    assert rel_norm(gI, gI64) <= 1e-4
    ctx.close()
