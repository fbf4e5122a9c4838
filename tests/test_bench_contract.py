"""bench.py's contract pieces that run without a GPU: the one-GPU refusal,
the matcher's algorithmic bytes, and the traffic figure read from the
committed counter summary."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_lk_algorithmic_bytes():
    shapes = [(1080 >> s, 1920 >> s) for s in range(4)]
    img = 8 * (1080 * 1920 + 540 * 960 + 270 * 480 + 135 * 240)
    assert bench.lk_algorithmic_bytes(shapes, [100, 10, 1, 0]) == img + 40 * 111


def test_traffic_from_counter_summary(tmp_path):
    t = bench.measured_traffic()
    assert t is not None and t > 0
    p = tmp_path / "summary.txt"
    p.write_text("header\ntraffic_bytes_per_launch = 1234\n")
    assert bench.measured_traffic(str(p)) == 1234
    assert bench.measured_traffic(str(tmp_path / "missing.txt")) is None


def test_valu_from_counter_summary(tmp_path):
    v = bench.measured_valu()
    assert v is not None and v > 0
    p = tmp_path / "summary.txt"
    p.write_text("header\nvalu_instr_per_level_solve = 43.5\n")
    assert bench.measured_valu(str(p)) == 43.5
    assert bench.measured_valu(str(tmp_path / "missing.txt")) is None
    # 10 particles at scale 0 of 4: 40 level solves in a 100 us launch
    tr = {"kernels": {"lk_track": {"avg_us": 100.0}}, "prof": {"frames": 1, "per_scale": [10, 0, 0, 0]},
          "shapes": [(1080 >> s, 1920 >> s) for s in range(4)]}
    r = bench.lk_roofline(tr)
    assert r["bound"] == "hbm" and 0 < r["frac"] < 1
    assert r["fp32"]["level_solves_per_launch"] == 40
    assert abs(r["fp32"]["achieved_tflops"] - 40 * bench.LK_FLOPS_PER_SOLVE / 100e-6 / 1e12) < 1e-3
    assert r["valu"]["instr_per_launch_est"] == int(v * 40)
    assert abs(r["valu"]["frac"] - v * 40 / 100e-6 / bench.VALU_ISSUE_PEAK) < 1e-3
