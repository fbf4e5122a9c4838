"""particle_container::add and compact(cpu) as standalone operations.

The oracle is pinned against plain Python loops over the reference's serial
order and against the reference's own assert after compaction
(particle_container.hpp:241: the particle at sparse(pos) has that position).
The GPU resolves the serial last-writer rule of the sparse image with a claim
(zero, then atomic maximum of the index) and must match the oracle exactly."""
import numpy as np
import pytest

import cuimg_amd as cu
import oracle

EMPTY = -1


def crafted(seed, n, nr, nc, dup_px=20, out_frac=0.05):
    """Particles crowded on few pixels (many per pixel), some dead, some outside
    the domain; a sparse image holding arbitrary earlier content."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n, oracle.PARTICLE_DTYPE)
    cells = rng.integers(0, nr * nc, max(1, min(dup_px, nr * nc)))
    pick = cells[rng.integers(0, len(cells), n)] if n > dup_px else rng.integers(0, nr * nc, n)
    p["pos_r"] = (pick // nc) + rng.random(n).astype(np.float32) * 0.99
    p["pos_c"] = (pick % nc) + rng.random(n).astype(np.float32) * 0.99
    out = rng.random(n) < out_frac
    p["pos_r"][out] = rng.choice([-3.5, nr + 2.25], int(out.sum()))
    p["age"] = rng.integers(0, 6, n)
    p["speed_r"] = rng.normal(0, 1, n)
    p["speed_c"] = rng.normal(0, 1, n)
    sparse = np.where(rng.random((nr, nc)) < 0.3, rng.integers(0, 1 << 20, (nr, nc)), EMPTY).astype(np.int32)
    return p, sparse


def py_compact(parts, sparse):
    sp = sparse.copy()
    nr, nc = sp.shape
    kept, matches = [], np.full(len(parts), 0xFFFFFFFF, np.uint32)
    for i in range(len(parts)):
        if parts["age"][i] != 0:
            j = len(kept)
            kept.append(parts[i])
            r, c = int(parts["pos_r"][i]), int(parts["pos_c"][i])
            if 0 <= r < nr and 0 <= c < nc:
                sp[r, c] = j
            matches[i] = j
    return np.array(kept, parts.dtype).reshape(-1), matches, sp


def py_add(parts, sparse, pts, cap):
    sp = sparse.copy()
    nr, nc = sp.shape
    out = list(parts)
    for r, c in pts:
        if len(out) >= cap:
            break
        q = np.zeros((), parts.dtype)
        q["age"] = 1
        q["pos_r"] = r
        q["pos_c"] = c
        if 0 <= r < nr and 0 <= c < nc:
            sp[r, c] = len(out)
        out.append(q)
    return np.array(out, parts.dtype).reshape(-1), sp


def rand_pts(seed, m, nr, nc):
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.integers(-2, nr + 2, m), rng.integers(-2, nc + 2, m)], 1).astype(np.int32)
    if m > 4:
        pts[m // 2] = pts[1]  # a repeated point: the later index wins
    return pts


# ------------------------------------------------------------------ oracle pins
@pytest.mark.parametrize("seed,n,nr,nc", [(1, 0, 6, 5), (2, 50, 8, 9), (3, 700, 16, 16), (4, 3000, 40, 30)])
def test_oracle_compact(orc, seed, n, nr, nc):
    p, sp = crafted(seed, n, nr, nc)
    got_p, got_m, got_s = orc.particles_compact(p, sp)
    want_p, want_m, want_s = py_compact(p, sp)
    assert np.array_equal(got_p, want_p)
    assert np.array_equal(got_m, want_m)
    assert np.array_equal(got_s, want_s)
    # the reference's assert after each write, at the end of the pass
    for j in range(len(got_p)):
        r, c = int(got_p["pos_r"][j]), int(got_p["pos_c"][j])
        if 0 <= r < nr and 0 <= c < nc:
            k = got_s[r, c]
            assert k >= j and int(got_p["pos_r"][k]) == r and int(got_p["pos_c"][k]) == c
    kept = got_m[got_m != 0xFFFFFFFF]
    assert np.array_equal(kept, np.arange(len(got_p)))  # monotone partial map


@pytest.mark.parametrize("seed,n,m,cap", [(1, 0, 30, 100), (2, 10, 50, 40), (3, 100, 0, 100), (4, 5, 200, 500)])
def test_oracle_add(orc, seed, n, m, cap):
    nr, nc = 20, 24
    p, sp = crafted(seed, n, nr, nc)
    pts = rand_pts(seed, m, nr, nc)
    got_p, got_s = orc.particles_add(p, sp, pts, cap)
    want_p, want_s = py_add(p, sp, pts, cap)
    assert np.array_equal(got_p, want_p)
    assert np.array_equal(got_s, want_s)


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


def same_parts(a, b):
    return all(np.array_equal(a[k], b[k]) for k in oracle.PARTICLE_DTYPE.names)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,nr,nc,dup", [(1, 0, 6, 5, 20), (2, 50, 8, 9, 20), (3, 700, 16, 16, 20),
                                              (4, 3000, 40, 30, 50), (5, 1025, 7, 300, 5),
                                              (6, 300000, 1080, 1920, 200000)])
def test_gpu_compact(ctx, orc, seed, n, nr, nc, dup):
    print(f"compact seed={seed} n={n} {nr}x{nc}", flush=True)
    p, sp = crafted(seed, n, nr, nc, dup_px=dup)
    want_p, want_m, want_s = orc.particles_compact(p, sp)
    got_p, got_m, got_s = cu.particles_compact(ctx, to_cu(p), sp)
    assert len(got_p) == len(want_p) and same_parts(got_p, want_p)
    assert np.array_equal(got_m, want_m)
    assert np.array_equal(got_s, want_s)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,m,cap,shape", [(1, 0, 30, 100, (20, 24)), (2, 10, 50, 40, (20, 24)),
                                                (3, 100, 0, 100, (20, 24)), (4, 5, 200, 500, (9, 9)),
                                                (5, 2000, 100000, 90000, (540, 960))])
def test_gpu_add(ctx, orc, seed, n, m, cap, shape):
    print(f"add seed={seed} n={n} m={m} cap={cap}", flush=True)
    nr, nc = shape
    p, sp = crafted(seed, n, nr, nc)
    pts = rand_pts(seed, m, nr, nc)
    want_p, want_s = orc.particles_add(p, sp, pts, cap)
    got_p, got_s = cu.particles_add(ctx, to_cu(p), sp, pts, cap)
    assert len(got_p) == len(want_p) and same_parts(got_p, want_p)
    assert np.array_equal(got_s, want_s)


@pytest.mark.gpu
def test_gpu_container_errors(ctx):
    p, sp = crafted(1, 10, 8, 8)
    with pytest.raises(cu.CuimgError):  # n above cap
        cu.particles_add(ctx, to_cu(p), sp, np.zeros((1, 2), np.int32), cap=5)
    src = cu.DeviceBuffer.from_array(ctx, to_cu(p))
    img = ctx.image(8, 8, cu.U32)
    n = cu.ctypes.c_uint32()
    rc = ctx.L.cuimg_hip_particles_compact(ctx.h, src.p, 10, src.p, src.p, cu.ctypes.byref(img.d),
                                           cu.ctypes.byref(n))
    assert rc != 0  # output overlapping the input
    src.free()
