"""merge_trajectories with the serial ascending-index semantics.

The oracle is pinned against an independent Python restatement of the serial
loop on crafted dense states (duplicates at one pixel, long removal chains).
The GPU resolves the serial loop in exact parallel rounds, then replays the
few actors they leave open serially in index order, in chunks of 64 lanes; a
Python model of that scheme is checked against the serial loop here, and the
GPU itself must reproduce the oracle exactly on states that take each of its
paths (open actors in LDS or in global memory; the short replay, and the full
sorted replay after the round limit)."""
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


def grid_state(seed, side, speed_sd=0.5):
    """One particle on every pixel of a side x side image, random order, ages
    and speeds: most particles are actors and long chains form."""
    rng = np.random.default_rng(seed)
    n = side * side
    p = np.zeros(n, oracle.PARTICLE_DTYPE)
    perm = rng.permutation(n)
    p["pos_r"] = (perm // side) + rng.random(n).astype(np.float32) * 0.999
    p["pos_c"] = (perm % side) + rng.random(n).astype(np.float32) * 0.999
    p["age"] = rng.integers(1, 30, n)
    p["speed_r"] = rng.normal(0, speed_sd, n)
    p["speed_c"] = rng.normal(0, speed_sd, n)
    sparse = np.full((side, side), -1, np.int32)
    sparse[p["pos_r"].astype(int), p["pos_c"].astype(int)] = np.arange(n)
    return p, sparse


def targets_of(parts, sparse):
    ages = parts["age"].astype(np.int64)
    nr, nc = sparse.shape
    n = len(parts)
    f32 = np.float32
    out = {}
    for i in range(n):
        if ages[i] == 0:
            continue
        r, c = int(parts["pos_r"][i]), int(parts["pos_c"][i])
        tl = []
        for dr, dc in C8:
            rr, cc = r + dr, c + dc
            if not (0 <= rr < nr and 0 <= cc < nc) or sparse[rr, cc] == -1 or sparse[rr, cc] >= n:
                continue
            b = int(sparse[rr, cc])
            d0 = f32(parts["speed_r"][i] - parts["speed_r"][b])
            d1 = f32(parts["speed_c"][i] - parts["speed_c"][b])
            if ages[b] > ages[i] + 2 and math.sqrt(f32(f32(d0 * d0) + f32(d1 * d1))) < 2.0:
                tl.append(b)
        if tl:
            out[i] = tl
    return out


def chain_state(n):
    """One row of n particles, index = column, each 3 frames older than its
    left neighbour: actor c targets c + 1 only, and the exact rounds resolve
    about one actor per round."""
    p = np.zeros(n, oracle.PARTICLE_DTYPE)
    p["pos_r"] = 0.5
    p["pos_c"] = np.arange(n) + 0.25
    p["age"] = 3 * np.arange(n) + 1
    sparse = np.arange(n, dtype=np.int32)[None, :].copy()
    return p, sparse


def rounds_merge(parts, sparse, switch=64, max_rounds=250, chunk=64):
    """Model of the GPU resolution: exact parallel rounds while more than
    `switch` actors are open (at most `max_rounds`), then the open actors in
    index order, chunk by chunk, seeing removals through zeroer (resolved
    actors and earlier chunks) or through the victims of earlier lanes."""
    T = targets_of(parts, sparse)
    BIG = 1 << 40
    zeroer = {}
    open_ = list(T)
    r = 0
    while open_ and (r == 0 or (len(open_) > switch and r < max_rounds)):
        r += 1
        minun = {}
        for i in open_:
            for b in T[i]:
                minun[b] = min(minun.get(b, BIG), i)
        nxt = []
        for i in open_:
            if minun.get(i, BIG) < i:
                nxt.append(i)
                continue
            if zeroer.get(i, BIG) < i:
                continue
            blocked, victim = False, -1
            for b in T[i]:
                if minun.get(b, BIG) < i:
                    blocked = True
                    break
                if zeroer.get(b, BIG) < i:
                    continue
                victim = b
                break
            if blocked:
                nxt.append(i)
            elif victim >= 0:
                zeroer[victim] = i
        open_ = nxt
    rest = sorted(open_)
    for c0 in range(0, len(rest), chunk):
        ids = rest[c0:c0 + chunk]
        zs = [zeroer.get(i, BIG) for i in ids]
        zt = [[zeroer.get(b, BIG) for b in T[i]] for i in ids]
        victims = []
        for l, i in enumerate(ids):
            v = -1
            if not (zs[l] < i or i in victims):
                for b, z in zip(T[i], zt[l]):
                    if not (z < i or b in victims):
                        v = b
                        break
            victims.append(v)
        for i, v in zip(ids, victims):
            if v >= 0:
                zeroer[v] = i
    ages = parts["age"].astype(np.int64).copy()
    for b in zeroer:
        ages[b] = 0
    return ages, r, len(rest)


@pytest.mark.parametrize("seed,n", [(1, 40), (2, 150), (3, 300), (4, 600), (5, 3000)])
def test_rounds_model_is_serial(seed, n):
    p, sparse = crafted_state(seed, n, (16, 16) if n <= 600 else (40, 40))
    ages, _, _ = rounds_merge(p, sparse)
    assert np.array_equal(ages, py_merge(p, sparse))


@pytest.mark.parametrize("side,switch,rounds", [(24, 64, 250), (40, 64, 250), (40, 0, 3), (40, 0, 1)])
def test_rounds_model_dense_grid(side, switch, rounds):
    """Rounds to completion, and the serial replay taking over after 1 or 3
    rounds with hundreds of open actors (many chunks)."""
    p, sparse = grid_state(side, side)
    ages, r, nrest = rounds_merge(p, sparse, switch, rounds)
    if rounds < 5:
        assert nrest > 128
    assert np.array_equal(ages, py_merge(p, sparse))


def test_rounds_model_long_chain():
    p, sparse = chain_state(400)
    ages, r, nrest = rounds_merge(p, sparse, 64, 100)
    assert r == 100 and nrest > 64
    assert np.array_equal(ages, py_merge(p, sparse))


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


@pytest.mark.gpu
@pytest.mark.parametrize("kind,size", [("grid", 40), ("grid", 100), ("chain", 700), ("chain", 5000)])
def test_gpu_merge_dense_states(orc, kind, size):
    """grid 40: about 1000 open actors after round 1 (LDS list); grid 100:
    about 6600 (beyond the 4096 LDS keys: global list); chain 700 and 5000:
    one actor per round, so the round limit is reached and the full sorted
    replay runs (sort in LDS, and in global memory)."""
    print(f"merge {kind} {size}", flush=True)
    ctx = cu.Context(0)
    p, sparse = grid_state(size, size) if kind == "grid" else chain_state(size)
    want = orc.merge(p, sparse)
    cp = np.zeros(len(p), cu.PARTICLE_DTYPE)
    for k in p.dtype.names:
        cp[k] = p[k]
    got = cu.particles_merge(ctx, cp, sparse)
    assert np.array_equal(got["age"], want["age"])
    # a second call on the same context sees clean scratch
    got2 = cu.particles_merge(ctx, cp, sparse)
    assert np.array_equal(got2["age"], want["age"])
    ctx.close()
