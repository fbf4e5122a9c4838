"""merge_trajectories with the serial ascending-index semantics.

The oracle is pinned against an independent Python restatement of the serial
loop on crafted dense states (duplicates at one pixel, long removal chains),
and the GPU's parallel round-based resolution must reproduce it exactly."""
import math

import numpy as np
import pytest

import cuimg_amd as cu
import oracle

C8 = [(-1, 1), (0, 1), (1, 1), (-1, 0), (1, 0), (-1, -1), (0, -1), (1, -1)]


def py_merge(parts, sparse):
    ages = parts["age"].astype(np.int64).copy()
    nr, nc = sparse.shape
    f32 = np.float32
    for i in range(len(parts)):
        if ages[i] == 0:
            continue
        r, c = int(parts["pos_r"][i]), int(parts["pos_c"][i])
        for dr, dc in C8:
            rr, cc = r + dr, c + dc
            if not (0 <= rr < nr and 0 <= cc < nc) or sparse[rr, cc] == -1:
                continue
            b = sparse[rr, cc]
            d0 = f32(parts["speed_r"][i] - parts["speed_r"][b])
            d1 = f32(parts["speed_c"][i] - parts["speed_c"][b])
            if ages[b] > ages[i] + 2 and math.sqrt(f32(f32(d0 * d0) + f32(d1 * d1))) < 2.0:
                ages[b] = 0
                break
    return ages


def crafted_state(seed, n, shape=(16, 16)):
    rng = np.random.default_rng(seed)
    nr, nc = shape
    p = np.zeros(n, oracle.PARTICLE_DTYPE)
    p["pos_r"] = rng.integers(0, nr, n) + rng.random(n).astype(np.float32) * 0.999
    p["pos_c"] = rng.integers(0, nc, n) + rng.random(n).astype(np.float32) * 0.999
    p["age"] = rng.integers(0, 14, n)
    p["speed_r"] = rng.normal(0, 1.0, n)
    p["speed_c"] = rng.normal(0, 1.0, n)
    sparse = np.full(shape, -1, np.int32)
    for i in range(n):  # last live particle of each pixel
        if p["age"][i]:
            sparse[int(p["pos_r"][i]), int(p["pos_c"][i])] = i
    return p, sparse


@pytest.mark.parametrize("seed,n", [(1, 40), (2, 150), (3, 300), (4, 600)])
def test_oracle_merge_is_serial(orc, seed, n):
    p, sparse = crafted_state(seed, n)
    got = orc.merge(p, sparse)
    assert np.array_equal(got["age"], py_merge(p, sparse))
    assert np.array_equal(got["pos_r"], p["pos_r"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n", [(1, 40), (2, 150), (3, 300), (4, 600), (5, 3000), (6, 0)])
def test_gpu_merge_matches_oracle(orc, seed, n):
    ctx = cu.Context(0)
    shape = (16, 16) if n <= 600 else (40, 40)
    p, sparse = crafted_state(seed, n, shape)
    want = orc.merge(p, sparse)
    cp = np.zeros(len(p), cu.PARTICLE_DTYPE)
    for k in p.dtype.names:
        cp[k] = p[k]
    got = cu.particles_merge(ctx, cp, sparse)
    assert np.array_equal(got["age"], want["age"])
    ctx.close()
