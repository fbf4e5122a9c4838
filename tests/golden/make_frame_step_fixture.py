"""Writes tests/golden/frame_step_320x240.json: the oracle tracker's state after
each of 10 frames of the seeded 320x240 translation video (3 scales, FAST
threshold 20), for the two matcher profiles (R: the reference's matcher on u8
frames; B: 7x7 / 5 steps on blurred FP32 frames) and the oracle's two
arithmetic readings ("raster": every product rounded, sums in loop order;
"gpu": fused multiply-adds and lane-group sums), and for the bc2s strategy
(C: integer arithmetic, one reading, the reference sample's detector
frequency of 10; its descriptors, flow and multiscale counts are stored too).

Stored per frame and scale: the dense vector's length, its live count, the
sum of ages, and SHA-256 digests of the dense vector (speed, position, age),
the sparse index image, the saliency and the matches map.  A regression pin:
any change of the frame step's results changes a digest.

Run:  python tests/golden/make_frame_step_fixture.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "frame_step_320x240.json")
SEED, NR, NC, NFRAMES, NSCALES, TH = 0xC0FFEE01, 240, 320, 10, 3, 20.0


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:32]


def particles_digest(p):
    """Digest of the fields a caller reads (the 2-byte pad is left out)."""
    q = np.zeros(len(p), [("speed_r", "<f4"), ("speed_c", "<f4"), ("pos_r", "<f4"), ("pos_c", "<f4"),
                          ("age", "<u2")])
    for k in q.dtype.names:
        q[k] = p[k]
    return digest(q)


def state(t, nscales):
    """Per scale: what the fixture stores of a tracker (oracle or GPU)."""
    out = []
    for s in range(nscales):
        p = t.particles(s)
        m = t.matches(s) if t.compact_has_run(s) else np.zeros(0, np.uint32)
        out.append({"n": int(len(p)), "live": int((p["age"] > 0).sum()), "age_sum": int(p["age"].sum()),
                    "particles": particles_digest(p), "sparse": digest(t.sparse(s).astype(np.int32)),
                    "saliency": digest(t.saliency(s)), "compact_has_run": bool(t.compact_has_run(s)),
                    "matches": digest(m.astype(np.uint32))})
    return out


def state_bc2s(t, nscales):
    """state() plus the bc2s descriptors, flow and multiscale counts."""
    out = state(t, nscales)
    for s in range(nscales):
        flow, msc = t.flow(s)
        out[s].update({"features": digest(t.features(s)), "flow": digest(flow.astype(np.float32)),
                       "msc": digest(msc.astype(np.uint32))})
    return out


C_DET_FREQ = 10


def video():
    import cuimg_amd as cu
    return cu.synth_video(SEED, NR, NC, NFRAMES, 0)


def main():
    sys.path.insert(0, ROOT)
    import oracle
    orc = oracle.Oracle()
    v = video()
    fixture = {"video": {"seed": SEED, "nrows": NR, "ncols": NC, "frames": NFRAMES, "motion": "translation",
                         "digest": digest(v)},
               "nscales": NSCALES, "fast_th": TH, "runs": {}}
    for profile in ("B", "R"):
        for arith in ("raster", "gpu"):
            ws = 7 if profile == "B" else 11
            lanes = oracle.gpu_sum_lanes(ws) if arith == "gpu" else 0
            t = orc.tracker(NR, NC, oracle.profile_params(profile, NSCALES, TH, sum_lanes=lanes))
            frames = []
            for f in range(NFRAMES):
                t.run(v[f])
                frames.append(state(t, NSCALES))
            fixture["runs"][f"{profile}/{arith}"] = frames
            print(profile, arith, [sc["live"] for sc in frames[-1]])
    t = orc.tracker(NR, NC, oracle.profile_params("C", NSCALES, TH, detector_frequency=C_DET_FREQ))
    frames = []
    for f in range(NFRAMES):
        t.run(v[f])
        frames.append(state_bc2s(t, NSCALES))
    fixture["runs"]["C/exact"] = frames
    print("C exact", [sc["live"] for sc in frames[-1]])
    with open(OUT, "w") as fh:
        json.dump(fixture, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
