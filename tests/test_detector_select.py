"""Detector mask and new-point selection as standalone operations.

The oracle's two functions are pinned against Python restatements written in
a different form (a per-pixel Chebyshev-distance test for the mask; a numpy
argmax per box, whose first-occurrence rule is the reference's first strict
maximum, for the selection).  The GPU entry points must then reproduce the
oracle exactly: both are integer work."""
import numpy as np
import pytest

import cuimg_amd as cu
import oracle


def rand_parts(rng, n, nr, nc, spread=0):
    p = np.zeros(n, oracle.PARTICLE_DTYPE)
    p["pos_r"] = rng.uniform(-1.5 - spread, nr + 0.5 + spread, n).astype(np.float32)
    p["pos_c"] = rng.uniform(-1.5 - spread, nc + 0.5 + spread, n).astype(np.float32)
    p["age"] = rng.integers(0, 5, n)  # age 0 (dead) particles still mask
    return p


def py_mask(parts, nr, nc):
    tr = parts["pos_r"].astype(np.int32)  # truncation toward zero
    tc = parts["pos_c"].astype(np.int32)
    rr, cc = np.mgrid[0:nr, 0:nc]
    m = np.full((nr, nc), 255, np.uint8)
    for r, c in zip(tr, tc):
        m[(np.abs(rr - r) <= 1) & (np.abs(cc - c) <= 1)] = 0
    return m


def rand_saliency(rng, nr, nc, zero_frac=0.6, levels=8):
    s = (rng.integers(0, levels, (nr, nc)) * (256 // levels)).astype(np.uint8)  # many ties
    s[rng.random((nr, nc)) < zero_frac] = 0
    return s


def rand_sparse(rng, nr, nc, frac):
    sp = np.full((nr, nc), -1, np.int32)
    hit = rng.random((nr, nc)) < frac
    sp[hit] = np.arange(int(hit.sum()), dtype=np.int32)
    return sp


def py_select(sal, sparse, box, th):
    nr, nc = sal.shape
    br, bc = int(np.float32(nr) / np.float32(box)), int(np.float32(nc) / np.float32(box))
    pts = []
    for i in range(br):
        for j in range(bc):
            r0, c0 = i * box + 1, j * box + 1
            win_sp = sparse[r0:r0 + box, c0:c0 + box]  # slicing clips to the domain
            if (win_sp != -1).any():
                continue
            win = np.zeros((box, box), np.int32)  # 0 outside the domain
            sub = sal[r0:r0 + box, c0:c0 + box]
            win[:sub.shape[0], :sub.shape[1]] = sub
            k = int(np.argmax(win))
            if np.float32(win.flat[k]) < np.float32(th):
                continue
            r, c = r0 + k // box, c0 + k % box
            if 3 <= r <= nr - 4 and 3 <= c <= nc - 4:
                pts.append((r, c))
    pts.sort()
    return np.array(pts, np.int32).reshape(-1, 2)


# ------------------------------------------------------------------ oracle pins
@pytest.mark.parametrize("seed,n,nr,nc", [(1, 0, 9, 7), (2, 30, 12, 17), (3, 200, 31, 23), (4, 64, 5, 5)])
def test_oracle_particle_mask(orc, seed, n, nr, nc):
    p = rand_parts(np.random.default_rng(seed), n, nr, nc)
    assert np.array_equal(orc.particle_mask(p, nr, nc), py_mask(p, nr, nc))


@pytest.mark.parametrize("box", [1, 2, 3, 5, 6, 9, 15])
@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_fast_select(orc, box, seed):
    rng = np.random.default_rng(seed * 100 + box)
    nr, nc = int(rng.integers(box, 70)), int(rng.integers(box, 90))
    sal = rand_saliency(rng, nr, nc)
    sp = rand_sparse(rng, nr, nc, 0.02)
    for th in (0.0, 32.0, 95.5, 300.0):
        got = orc.fast_select(sal, sp, box, th)
        want = py_select(sal, sp, box, th)
        assert np.array_equal(got, want), (box, th)


def test_oracle_fast_select_plateau_takes_first(orc):
    """A box of equal values yields its first pixel; a later equal value never
    replaces an earlier maximum (strict comparison)."""
    sal = np.zeros((20, 20), np.uint8)
    sal[5:9, 5:9] = 50          # box (1,1) of size 4 covers rows/cols 5..8: all 50
    sal[13, 14] = 90            # box (3,3) covers rows/cols 13..16
    sal[14, 13] = 90
    sp = np.full((20, 20), -1, np.int32)
    got = orc.fast_select(sal, sp, 4, 10.0)
    assert [tuple(x) for x in got] == [(5, 5), (13, 14)]
    sp[15, 15] = 0              # a particle in box (3,3) removes it
    got = orc.fast_select(sal, sp, 4, 10.0)
    assert [tuple(x) for x in got] == [(5, 5)]


# --------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def ctx():
    c = cu.Context(0)
    yield c
    c.close()


def to_cu(p):
    out = np.zeros(len(p), cu.PARTICLE_DTYPE)
    for k in p.dtype.names:
        out[k] = p[k]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,nr,nc", [(1, 0, 9, 7), (2, 30, 12, 17), (3, 2000, 131, 97),
                                          (4, 64, 5, 5), (5, 300000, 1080, 1920), (6, 5000, 1, 700)])
def test_gpu_particle_mask(ctx, orc, seed, n, nr, nc):
    print(f"particle_mask seed={seed} n={n} {nr}x{nc}", flush=True)
    p = rand_parts(np.random.default_rng(seed), n, nr, nc, spread=2)
    want = orc.particle_mask(p, nr, nc)
    got = cu.particle_mask(ctx, to_cu(p), nr, nc)
    assert np.array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("box", [1, 2, 3, 4, 6, 8, 11, 15])
@pytest.mark.parametrize("shape", [(37, 53), (240, 320), (3, 40), (1080, 1920)])
def test_gpu_fast_select(ctx, orc, box, shape):
    nr, nc = shape
    print(f"fast_select box={box} {nr}x{nc}", flush=True)
    rng = np.random.default_rng(box * 7919 + nr)
    sal = rand_saliency(rng, nr, nc)
    sp = rand_sparse(rng, nr, nc, 0.01)
    for th in (0.0, 40.0, 96.25):
        want = orc.fast_select(sal, sp, box, th)
        got, n = cu.fast_select(ctx, sal, sp, box, th)
        assert n == len(want)
        assert np.array_equal(got, want), (box, th)


@pytest.mark.gpu
def test_gpu_fast_select_cap_and_errors(ctx, orc):
    rng = np.random.default_rng(77)
    sal = rand_saliency(rng, 200, 300, zero_frac=0.1)
    sp = np.full((200, 300), -1, np.int32)
    want = orc.fast_select(sal, sp, 5, 20.0)
    assert len(want) > 100
    got, n = cu.fast_select(ctx, sal, sp, 5, 20.0, cap=100)
    assert n == len(want) and np.array_equal(got, want[:100])
    got, n = cu.fast_select(ctx, sal[:4, :4], sp[:4, :4], 5, 20.0)  # no whole box
    assert n == 0 and len(got) == 0
    with pytest.raises(cu.CuimgError):
        cu.fast_select(ctx, sal, sp, 16, 20.0)
    with pytest.raises(cu.CuimgError):
        cu.fast_select(ctx, sal, sp, 0, 20.0)
