"""Frame-by-frame comparison of a GPU tracker with the oracle tracker (lockstep).

Before every frame both trackers hold the same particle vectors (the GPU one
receives the oracle's), both run the same frame, and the frame's outputs are
compared:
  * pyramid levels: bit-exact;
  * matcher decisions: accept flags, |translation difference| <= 0.01 px, and
    (with the oracle in the GPU's arithmetic, the default here) bit-for-bit
    equality of translation and distance;
  * when every decision agrees and the truncated new positions agree, the rest
    of the frame is integer work and must match exactly: saliency, the dense
    particle vector (order, ages), matches and the sparse index image
    ("exact" frame);
  * otherwise ("partial" frame) everything out of reach of the disagreeing
    particles is still compared exactly: the saliency and the sparse image away
    from them, the new points of boxes they cannot touch (positions and raster
    order), and every other old particle's fate (kept or removed, age,
    position, speed, its matches entry).
"""
import numpy as np

TOL_PX = 0.01
EMPTY = 0xFFFFFFFF
# Chebyshev radius around a disagreeing particle's old and new pixels within
# which results may differ: the detector mask (1), the merge neighbourhood (1)
# and a buddy of a buddy.
REACH = 3


def to_cu(parts):
    import cuimg_amd as cu
    out = np.zeros(len(parts), cu.PARTICLE_DTYPE)
    for k in parts.dtype.names:
        out[k] = parts[k]
    return out


def new_stats():
    return dict(decisions=0, decision_mismatch=0, tracked=0, within=0, bit_exact=0, frames_exact=0,
                frames_partial=0, partial_checked=0, partial_skipped=0)


def _trunc(p):
    return np.stack([p["pos_r"].astype(np.int32), p["pos_c"].astype(np.int32)], 1)


def _moved(pre, tr):
    r = (pre["pos_r"] + tr[:, 0]).astype(np.float32)
    c = (pre["pos_c"] + tr[:, 1]).astype(np.float32)
    return np.stack([r.astype(np.int32), c.astype(np.int32)], 1)


def _near_mask(shape, pts, radius):
    """Pixels within Chebyshev distance `radius` of any point (clipped)."""
    nr, nc = shape
    m = np.zeros((nr + 2 * radius + 1, nc + 2 * radius + 1), bool)
    if len(pts):
        p = np.asarray(pts, np.int64)
        keep = (p[:, 0] >= -radius) & (p[:, 0] < nr + radius) & (p[:, 1] >= -radius) & (p[:, 1] < nc + radius)
        p = p[keep]
        for dr in range(-radius, radius + 1):
            for dc in range(-radius, radius + 1):
                m[p[:, 0] + dr + radius, p[:, 1] + dc + radius] = True
    return m[radius:radius + nr, radius:radius + nc]


def _pixel_hit(mask, pts):
    p = np.asarray(pts, np.int64).reshape(-1, 2)
    nr, nc = mask.shape
    inside = (p[:, 0] >= 0) & (p[:, 0] < nr) & (p[:, 1] >= 0) & (p[:, 1] < nc)
    out = np.zeros(len(p), bool)
    out[inside] = mask[p[inside, 0], p[inside, 1]]
    return out


def _partial_checks(o, g, s, pre, pre_g, acc_o, tr_o, acc_g, tr_g, box, stats):
    """Exact checks away from the disagreeing particles of scale s."""
    nr, nc = o.dims(s)
    n_pre = len(pre)
    po, pg = o.particles(s), g.particles(s)
    new_o, new_g = _moved(pre, tr_o), _moved(pre_g, tr_g)
    both = (acc_o == 1) & (acc_g == 1)
    far_tr = np.abs(tr_o - tr_g).max(axis=1, initial=0) > TOL_PX
    old, old_g = _trunc(pre), _trunc(pre_g)
    differ = (acc_o != acc_g) | np.any(old != old_g, axis=1) | (both & (np.any(new_o != new_g, axis=1) | far_tr))
    bad_pts = np.concatenate([old[differ], old_g[differ], new_o[differ & (acc_o == 1)],
                              new_g[differ & (acc_g == 1)]])
    near = _near_mask((nr, nc), bad_pts, REACH)

    # saliency and sparse-image emptiness away from the disagreeing particles
    so, sg = o.saliency(s), g.saliency(s)
    assert np.array_equal(so[~near], sg[~near]), f"saliency away from disagreeing particles, scale {s}"
    spo, spg = o.sparse(s), g.sparse(s)
    assert np.array_equal(spo[~near] == -1, spg[~near] == -1), f"sparse occupancy, scale {s}"

    # fate of the other old particles: through matches after a compaction,
    # index for index otherwise
    assert o.compact_has_run(s) == g.compact_has_run(s)
    if o.compact_has_run(s):
        mo, mg = o.matches(s), g.matches(s)
        assert len(mo) >= n_pre and len(mg) >= n_pre
        mo, mg = mo[:n_pre].astype(np.int64), mg[:n_pre].astype(np.int64)
    else:
        assert len(po) == len(pg) == n_pre
        mo = mg = np.arange(n_pre, dtype=np.int64)
    clean = ~differ & ~_pixel_hit(near, old) & ~_pixel_hit(near, new_o)
    ci = np.nonzero(clean)[0]
    assert np.array_equal(mo[ci] == EMPTY, mg[ci] == EMPTY), f"removed particles, scale {s}"
    kept = ci[mo[ci] != EMPTY]
    a, b = po[mo[kept]], pg[mg[kept]]
    assert np.array_equal(a["age"], b["age"]), f"ages of kept particles, scale {s}"
    for f in ("pos_r", "pos_c", "speed_r", "speed_c"):
        assert np.abs(a[f] - b[f]).max(initial=0) <= TOL_PX, f"{f} of kept particles, scale {s}"
    # the sparse index of a kept clean particle's pixel names the same particle
    t = _trunc(a)
    ins = (t[:, 0] >= 0) & (t[:, 0] < nr) & (t[:, 1] >= 0) & (t[:, 1] < nc)
    t = t[ins]
    io, ig = spo[t[:, 0], t[:, 1]], spg[t[:, 0], t[:, 1]]
    both = (io >= 0) & (ig >= 0)
    assert np.array_equal(io >= 0, ig >= 0), f"sparse entries of kept particles, scale {s}"
    assert np.array_equal(_trunc(po[io[both]]), _trunc(pg[ig[both]])), f"sparse targets, scale {s}"
    assert np.array_equal(po[io[both]]["age"], pg[ig[both]]["age"]), f"sparse target ages, scale {s}"

    # new points (age 1 after a detection) of boxes the disagreeing particles
    # cannot reach: same points in the same raster order
    if o.compact_has_run(s):
        no, ng = _trunc(po[po["age"] == 1]), _trunc(pg[pg["age"] == 1])
        bnear = np.zeros(((nr + box - 1) // box + 1, (nc + box - 1) // box + 1), bool)
        rr, cc = np.nonzero(near)
        bnear[(rr - 1).clip(0) // box, (cc - 1).clip(0) // box] = True

        def far(p):
            return p[~bnear[(p[:, 0] - 1).clip(0) // box, (p[:, 1] - 1).clip(0) // box]]

        assert np.array_equal(far(no), far(ng)), f"new points away from disagreeing particles, scale {s}"
    stats["partial_checked"] += int(len(ci))
    stats["partial_skipped"] += int(n_pre - len(ci))


def compare_frame(o, g, L, stats, pre, box=3, pre_g=None):
    """Compares one lockstep frame; pre[s] is the particle vector both trackers
    held at scale s before it (pre_g[s] the GPU tracker's, when it differs:
    only the checks of this module use that)."""
    pre_g = pre if pre_g is None else pre_g
    frame_exact = True
    dec = []
    for s in range(L):
        acc_o, tr_o, dist_o = o.last_decisions(s)
        acc_g, tr_g, dist_g = g.decisions(s)
        assert len(acc_o) == len(acc_g) == len(pre[s]), (s, len(acc_o), len(acc_g), len(pre[s]))
        dec.append((acc_o, tr_o, acc_g, tr_g))
        stats["bit_exact"] += int(((acc_o == acc_g) & np.all(tr_o.view(np.uint32) == tr_g.view(np.uint32), axis=1)
                                   & (dist_o.view(np.uint32) == dist_g.view(np.uint32))).sum())
        same = acc_o == acc_g
        stats["decisions"] += len(acc_o)
        stats["decision_mismatch"] += int((~same).sum())
        both = same & (acc_o == 1)
        d = np.abs(tr_o[both] - tr_g[both]).max(axis=1) if both.any() else np.zeros(0)
        stats["tracked"] += int(both.sum())
        stats["within"] += int((d <= TOL_PX).sum())
        if (~same).any() or (d > TOL_PX).any():
            frame_exact = False
    for s in range(L):
        assert np.array_equal(o.level(s), g.level(s)), f"pyramid level {s}"
    if frame_exact:
        for s in range(L):
            po, pg = o.particles(s), g.particles(s)
            if len(po) != len(pg) or not np.array_equal(_trunc(po), _trunc(pg)):
                frame_exact = False
                break
    if not frame_exact:
        for s in range(L):
            _partial_checks(o, g, s, pre[s], pre_g[s], *dec[s], box, stats)
        stats["frames_partial"] += 1
        return
    for s in range(L):
        assert np.array_equal(o.saliency(s), g.saliency(s)), f"saliency {s}"
        po, pg = o.particles(s), g.particles(s)
        assert np.array_equal(po["age"], pg["age"]), f"ages {s}"
        assert np.abs(po["pos_r"] - pg["pos_r"]).max(initial=0) <= TOL_PX
        assert np.abs(po["pos_c"] - pg["pos_c"]).max(initial=0) <= TOL_PX
        assert o.compact_has_run(s) == g.compact_has_run(s)
        if o.compact_has_run(s):
            assert np.array_equal(o.matches(s), g.matches(s)), f"matches {s}"
        assert np.array_equal(o.sparse(s), g.sparse(s)), f"sparse {s}"
    stats["frames_exact"] += 1


def run_lockstep(ctx, orc, video, params_o, params_g, nframes, log=True, sum_order="gpu"):
    """sum_order "gpu": the oracle adds the matcher's window sums in the GPU's
    lane-group order (decisions can then agree bit for bit); "raster": in the
    reference's loop order (they agree to FP32 reassociation)."""
    import cuimg_amd as cu
    import oracle
    nr, nc = video.shape[1:]
    L = params_o.nscales
    params_o = type(params_o).from_buffer_copy(params_o)
    params_o.sum_lanes = oracle.gpu_sum_lanes(params_o.lk_ws) if sum_order == "gpu" else 0
    o = orc.tracker(nr, nc, params_o)
    g = cu.Tracker(ctx, nr, nc, params_g)
    stats = new_stats()
    for f in range(nframes):
        if log:
            print(f"frame {f}", flush=True)
        pre = [o.particles(s) for s in range(L)]
        for s in range(L):
            g.set_particles(s, to_cu(pre[s]))
        o.run(video[f])
        g.run(video[f])
        assert g.frame_count == f + 1
        compare_frame(o, g, L, stats, pre, params_o.box_size)
    g.close()
    return stats
