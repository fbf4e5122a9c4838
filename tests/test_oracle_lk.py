"""Pins the oracle's Lucas-Kanade matcher.

Mathematical checks: zero motion gives exactly zero, integer and sub-pixel
translations of smooth textures are recovered, degenerate (flat / 1-D) patches
fail with the reference's failure code, leaving the domain fails, the Sobel/8
gradient matches its closed form.  An independent FP32 restatement in numpy
(scalar float32 arithmetic in the same operation order) checks the window
bookkeeping (selection with ties, the error's index and counter) bit-exactly.
"""
import numpy as np
import pytest

import oracle

f32 = np.float32


def smooth_texture(rng, shape, sigma=2.0):
    from scipy.ndimage import gaussian_filter
    t = gaussian_filter(rng.standard_normal(shape), sigma)
    t = (t - t.min()) / (t.max() - t.min()) * 255.0
    return t.astype(np.float32)


def test_sobel_closed_form(orc):
    rng = np.random.default_rng(0)
    P = rng.integers(0, 256, (20, 30)).astype(np.float32)
    G = orc.sobel(P, True)
    Pi = P.astype(np.int64)
    gr = (Pi[2:, :-2] + 2 * Pi[2:, 1:-1] + Pi[2:, 2:] - Pi[:-2, :-2] - 2 * Pi[:-2, 1:-1] - Pi[:-2, 2:]) / 8
    gc = (Pi[:-2, 2:] + 2 * Pi[1:-1, 2:] + Pi[2:, 2:] - Pi[:-2, :-2] - 2 * Pi[1:-1, :-2] - Pi[2:, :-2]) / 8
    assert np.array_equal(G[1:-1, 1:-1, 0], gr.astype(np.float32))
    assert np.array_equal(G[1:-1, 1:-1, 1], gc.astype(np.float32))
    assert (G[0] == 0).all() and (G[-1] == 0).all() and (G[:, 0] == 0).all() and (G[:, -1] == 0).all()


@pytest.mark.parametrize("profile", ["R", "B"])
def test_zero_motion_is_exact(orc, profile):
    rng = np.random.default_rng(1)
    A = rng.integers(0, 256, (40, 50)).astype(np.float32)
    G = orc.sobel(A, True)
    kw = dict(ws=11, max_iters=30, top_n=25, template_u8=1) if profile == "R" else dict(ws=7, max_iters=5, top_n=49)
    ws = kw["ws"]
    hws, n = ws // 2, ws * ws
    nsel = min(kw["top_n"], n)
    for p in [(20, 25), (10, 12)]:
        tr, dist = orc.lk_match(A, A, G, (float(p[0]), float(p[1])), (0.0, 0.0), **kw)
        assert tr[0] == 0.0 and tr[1] == 0.0
        # the error pairs window pixel (r,c) with template index (r+hws)*hws + (c+hws)
        # (tracking2/pyrlk_tracker.hpp:325): non-zero even without motion
        err = 0.0
        for r in range(-hws, hws + 1):
            for c in range(-hws, hws + 1):
                idx = (r + hws) * hws + (c + hws)
                rr, cc = idx // ws - hws, idx % ws - hws
                err += abs(float(A[p[0] + rr, p[1] + cc]) - float(A[p[0] + r, p[1] + c]))
        assert abs(dist - err / (nsel + n)) <= 1e-4 * max(1.0, dist)


@pytest.mark.parametrize("shift", [(2, -1), (-1, 1), (0, 2)])
def test_integer_translation_recovered(orc, shift):
    rng = np.random.default_rng(2)
    A = smooth_texture(rng, (60, 60))
    B = np.roll(A, shift, axis=(0, 1))
    G = orc.sobel(A, False)
    tr, _ = orc.lk_match(A, B, G, (30.0, 30.0), (0.0, 0.0), ws=11, max_iters=50, eps=1e-4, top_n=121)
    assert np.abs(tr - np.array(shift, np.float32)).max() < 1e-2, tr


@pytest.mark.parametrize("shift", [(0.4, -0.3), (1.25, 0.75)])
def test_subpixel_translation_recovered(orc, shift):
    rng = np.random.default_rng(3)
    A = smooth_texture(rng, (80, 80), 3.0)
    from scipy.ndimage import map_coordinates
    rr, cc = np.mgrid[0:80, 0:80].astype(np.float64)
    B = map_coordinates(A.astype(np.float64), [rr - shift[0], cc - shift[1]], order=1, mode="nearest").astype(np.float32)
    G = orc.sobel(A, False)
    tr, _ = orc.lk_match(A, B, G, (40.0, 40.0), (0.0, 0.0), ws=11, max_iters=50, eps=1e-4, top_n=121)
    assert np.abs(tr - np.array(shift, np.float32)).max() < 0.05, tr


def test_degenerate_patches_fail(orc):
    flat = np.full((40, 40), 77, np.float32)
    tr, dist = orc.lk_match(flat, flat, orc.sobel(flat, True), (20.0, 20.0), (0.0, 0.0))
    assert tuple(tr) == (-1.0, -1.0) and dist == 1000.0
    stripes = np.tile((np.arange(40) * 7 % 256).astype(np.float32), (40, 1))  # varies along columns only
    tr, dist = orc.lk_match(stripes, stripes, orc.sobel(stripes, True), (20.0, 20.0), (0.0, 0.0))
    assert tuple(tr) == (-1.0, -1.0) and dist == 1000.0


def test_leaving_domain_fails(orc):
    rng = np.random.default_rng(4)
    A = smooth_texture(rng, (40, 40))
    B = np.roll(A, (-6, 0), axis=(0, 1))  # content moved up past the border band
    G = orc.sobel(A, False)
    tr, dist = orc.lk_match(A, B, G, (5.0, 20.0), (0.0, 0.0), ws=7, max_iters=20)
    assert tuple(tr) == (-1.0, -1.0) and dist == 1000.0


# ----------------------------------------------------------- numpy restatement
def np_li(img, qr, qc, fused=False):
    nr, nc = img.shape
    xr, xc = int(qr), int(qc)
    a0, a1 = f32(f32(qr) - f32(xr)), f32(f32(qc) - f32(xc))
    g = lambda r, c: img[min(max(r, 0), nr - 1), min(max(c, 0), nc - 1)]
    b0, b1 = f32(f32(1) - a0), f32(f32(1) - a1)
    s = f32(f32(b0 * b1) * g(xr, xc))
    terms = ((f32(a0 * b1), g(xr + 1, xc)), (f32(b0 * a1), g(xr, xc + 1)), (f32(a0 * a1), g(xr + 1, xc + 1)))
    for w, v in terms:
        s = f32_fma(w, v, s) if fused else f32(s + f32(w * v))
    return s


def f32_fma(a, b, c):
    """fma in float32: a*b + c exactly (fractions), rounded once to nearest
    even."""
    from fractions import Fraction
    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    x = np.float32(float(v))
    best = None
    for y in (np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))):
        d = abs(Fraction(float(y)) - v)
        even = (int(np.float32(y).view(np.uint32)) & 1) == 0
        if best is None or d < best[0] or (d == best[0] and even):
            best = (d, y)
    return np.float32(best[1])


class Sum:
    """A window sum in raster order (lanes 0) or in the lane-group order: for
    16 lanes pixel (r, c) of the window goes to lane 8 (r % 2) + c, for 64
    lanes pixel i to lane i % 64; lanes are added pairwise in rows of 16, rows
    as (r3 + r2) + (r1 + r0)."""

    def __init__(self, lanes, ws):
        self.lanes, self.ws = lanes, ws
        self.acc = [f32(0)] * max(lanes, 1)

    def lane(self, i):
        if not self.lanes:
            return 0
        if self.lanes == 16:
            return ((i // self.ws) % 2) * 8 + i % self.ws
        return i % self.lanes

    def fma(self, i, a, b):
        if self.lanes:
            self.acc[self.lane(i)] = f32_fma(a, b, self.acc[self.lane(i)])
        else:
            self.acc[0] = f32(self.acc[0] + f32(a * b))

    def add(self, i, x):
        k = self.lane(i)
        self.acc[k] = f32(self.acc[k] + x)

    def total(self):
        if not self.lanes:
            return self.acc[0]

        def tree(x):
            while len(x) > 1:
                x = [f32(x[2 * j] + x[2 * j + 1]) for j in range(len(x) // 2)]
            return x[0]
        rows = [tree(self.acc[16 * r:16 * r + 16]) for r in range(self.lanes // 16)]
        return rows[0] if len(rows) == 1 else f32(f32(rows[3] + rows[2]) + f32(rows[1] + rows[0]))


def np_match(A, B, G, p, pred, ws, max_iters, eps, top_n, template_u8, lanes=0):
    nr, nc = A.shape
    hws, n = ws // 2, ws * ws
    fz = lanes != 0  # the GPU's arithmetic: fused bilinear blends with the lane-group sums
    p = (f32(p[0]), f32(p[1]))
    prediction = (f32(p[0] + f32(pred[0])), f32(p[1] + f32(pred[1])))
    v = list(prediction)
    gs, as_, valid = [], [], []
    for r in range(-hws, hws + 1):
        for c in range(-hws, hws + 1):
            qr, qc = f32(p[0] + f32(r)), f32(p[1] + f32(c))
            if 0 <= int(qr) < nr and 0 <= int(qc) < nc:
                g = (np_li(G[:, :, 0], qr, qc, fz), np_li(G[:, :, 1], qr, qc, fz))
                a = np_li(A, qr, qc, fz)
                if template_u8:
                    a = f32(np.uint8(int(a) & 0xFF))
                gs.append(g), as_.append(a), valid.append(True)
            else:
                gs.append((f32(0), f32(0))), as_.append(f32(0)), valid.append(False)
    mag = [np.sqrt(f32(f32(g[0] * g[0]) + f32(g[1] * g[1]))) for g in gs]
    order = sorted([i for i in range(n) if valid[i]], key=lambda i: (-mag[i], i))
    sel = set(order[:top_n])
    s00, s01, s11 = Sum(lanes, ws), Sum(lanes, ws), Sum(lanes, ws)
    for i in range(n):
        if i in sel:
            gx, gy = gs[i]
            s00.fma(i, gx, gx)
            s01.fma(i, gx, gy)
            s11.fma(i, gy, gy)
    G00, G01, G11 = s00.total(), s01.total(), s11.total()
    h, q = f32(f32(0.5) * f32(G00 + G11)), f32(f32(0.5) * f32(G00 - G11))
    rr = np.sqrt(f32(f32(q * q) + f32(G01 * G01)))
    if min(abs(f32(h + rr)), abs(f32(h - rr))) < f32(1e-4):
        return np.array([-1, -1], np.float32), f32(1000)
    det = f32(f32(G00 * G11) - f32(G01 * G01))
    inv = f32(f32(1) / det)
    I00, I01, I11 = f32(G11 * inv), f32(-G01 * inv), f32(G00 * inv)
    nk = (f32(1), f32(1))
    k, cpt = 0, len(sel)
    while k < max_iters and np.sqrt(f32(f32(nk[0] * nk[0]) + f32(nk[1] * nk[1]))) >= f32(eps):
        sb0, sb1 = Sum(lanes, ws), Sum(lanes, ws)
        cpt = 0
        i = 0
        for r in range(-hws, hws + 1):
            for c in range(-hws, hws + 1):
                if i in sel:
                    dt = f32(as_[i] - np_li(B, f32(v[0] + f32(r)), f32(v[1] + f32(c)), fz))
                    sb0.fma(i, gs[i][0], dt)
                    sb1.fma(i, gs[i][1], dt)
                    cpt += 1
                i += 1
        b0, b1 = sb0.total(), sb1.total()
        nk = (f32(f32(I00 * b0) + f32(I01 * b1)), f32(f32(I01 * b0) + f32(I11 * b1)))
        v = [f32(v[0] + nk[0]), f32(v[1] + nk[1])]
        d0, d1 = f32(prediction[0] - v[0]), f32(prediction[1] - v[1])
        if np.sqrt(f32(f32(d0 * d0) + f32(d1 * d1))) > ws * 4:
            return np.array([-1, -1], np.float32), f32(1000)
        if not (3 <= int(v[0]) <= nr - 4 and 3 <= int(v[1]) <= nc - 4):
            return np.array([-1, -1], np.float32), f32(1000)
        k += 1
    se = Sum(lanes, ws)
    w = 0
    for r in range(-hws, hws + 1):
        for c in range(-hws, hws + 1):
            idx = (r + hws) * hws + (c + hws)
            se.add(w, abs(f32(as_[idx] - np_li(B, f32(v[0] + f32(r)), f32(v[1] + f32(c)), fz))))
            cpt += 1
            w += 1
    return np.array([f32(v[0] - p[0]), f32(v[1] - p[1])], np.float32), f32(se.total() / f32(cpt))


@pytest.mark.parametrize("lanes", [0, "gpu"])
@pytest.mark.parametrize("profile", ["R", "B"])
def test_matches_independent_restatement(orc, profile, lanes):
    """Bit-exact agreement with the numpy restatement, with window sums in
    raster order and in the GPU's lane-group order."""
    rng = np.random.default_rng(10 if profile == "R" else 11)
    kw = (dict(ws=11, max_iters=30, eps=0.1, top_n=25, template_u8=1) if profile == "R"
          else dict(ws=7, max_iters=5, eps=0.1, top_n=49, template_u8=0))
    kw["lanes"] = oracle.gpu_sum_lanes(kw["ws"]) if lanes == "gpu" else 0
    for trial in range(12):
        if profile == "R":
            A = rng.integers(0, 256, (36, 40)).astype(np.float32)
            A[::3, ::2] = 128  # plateaus: equal gradients, exercising the tie rule
            G = orc.sobel(A, True)
            B = np.clip(np.roll(A, (1, -1), axis=(0, 1)) + rng.integers(-3, 4, A.shape), 0, 255).astype(np.float32)
        else:
            A = smooth_texture(rng, (36, 40))
            G = orc.sobel(A, False)
            B = np.roll(A, (1, 0), axis=(0, 1)) + rng.normal(0, 1, A.shape).astype(np.float32)
        p = (float(rng.uniform(2, 33)), float(rng.uniform(2, 37)))
        pred = (float(rng.uniform(-1.5, 1.5)), float(rng.uniform(-1.5, 1.5)))
        okw = {k: v for k, v in kw.items() if k != "lanes"}
        tr, dist = orc.lk_match(A, B, G, p, pred, sum_lanes=kw["lanes"], **okw)
        tr2, dist2 = np_match(A, B, G, p, pred, **kw)
        assert np.array_equal(tr, tr2), (trial, tr, tr2)
        assert np.float32(dist) == dist2, (trial, dist, dist2)


def _texture(seed, nr, nc):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    t = gaussian_filter(rng.normal(0, 1, (nr, nc)), 2.0)
    t = (t - t.min()) / (t.max() - t.min()) * 255.0
    return np.round(t).astype(np.uint8)


def test_oracle_recursion_recovers_large_translation(orc):
    """recursive_prediction (predictions.h:60-74, get_flow_at
    pyrlk_tracker.hpp:1025-1063): a shift of (+12, -18) px is far beyond one
    7x7 level; the 4-level coarse-to-fine recursion (prediction 2 x the coarser
    flow) recovers it for the particles detected on the first frame, while the
    same frames on one level do not."""
    canvas = _texture(11, 400, 500)
    dy, dx = 12, -18
    f0 = np.ascontiguousarray(canvas[60:300, 60:380])
    f1 = np.ascontiguousarray(canvas[60 - dy:300 - dy, 60 - dx:380 - dx])  # content moves by (dy, dx)
    t = orc.tracker(240, 320, oracle.profile_params("B", 4, 8.0))
    t.run(f0)
    t.run(f1)
    p = t.particles(0)
    moved = p[p["age"] >= 2]
    assert len(moved) > 100
    sr, sc = moved["speed_r"].astype(np.float64), moved["speed_c"].astype(np.float64)
    assert abs(np.median(sr) - dy) < 0.05 and abs(np.median(sc) - dx) < 0.05
    close = (np.abs(sr - dy) < 0.5) & (np.abs(sc - dx) < 0.5)
    assert close.mean() > 0.5  # the rest see content that entered the frame
    # the same frames on one level: the shift is out of the matcher's reach
    t1 = orc.tracker(240, 320, oracle.profile_params("B", 1, 8.0))
    t1.run(f0)
    t1.run(f1)
    q = t1.particles(0)
    q = q[q["age"] >= 2]
    assert abs(np.median(q["speed_r"]) - dy) > 5 and abs(np.median(q["speed_c"]) - dx) > 5


def _static_invariants(tracker, nscales, nframes, frame):
    tracker.run(frame)
    first = [tracker.particles(s).copy() for s in range(nscales)]
    assert sum(len(p) for p in first) > 100
    for k in range(1, nframes):
        tracker.run(frame)
        for s in range(nscales):
            p = tracker.particles(s)
            n0 = len(first[s])
            old, new = p[:n0], p[n0:]
            assert len(p) >= n0, (k, s)
            assert np.array_equal(old["pos_r"], first[s]["pos_r"]) and np.array_equal(old["pos_c"], first[s]["pos_c"])
            assert (p["speed_r"] == 0).all() and (p["speed_c"] == 0).all()
            assert (old["age"] == k + 1).all(), (k, s)
            # a box whose first maximum sat in the 3-pixel border can get a
            # new in-domain maximum once a neighbour's mask zeroes it: rare
            assert len(new) <= max(2, n0 // 50) and (new["age"] <= k).all(), (k, s, len(new))
            sp = tracker.sparse(s)
            idx = sp[p["pos_r"].astype(np.int64), p["pos_c"].astype(np.int64)]
            assert np.array_equal(idx, np.arange(len(p))), (k, s)
            assert (sp >= 0).sum() == len(p)


def test_oracle_static_video_known_answer(orc):
    """Frame step + container (tracker.hpp:12-112, particle_container move /
    compact / add) on identical frames: every particle matches in place (zero
    translation, accepted, age + 1 per frame) and keeps its place in the dense
    vector, the detector adds almost nothing after the first frame (occupied
    boxes are skipped), nothing merges (the original particles share one age),
    and the sparse image indexes exactly the dense vector."""
    frame = _texture(5, 160, 224)
    _static_invariants(orc.tracker(160, 224, oracle.profile_params("B", 3, 20.0)), 3, 5, frame)
