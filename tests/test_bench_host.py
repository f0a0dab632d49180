"""Host-side checks of bench.py: the frame schedule, the profile readers and
the one-GPU rule. No GPU work."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("steps", [1, 5, 20, 24, 60, 119, 120, 250])
def test_schedule_covers_the_path(steps):
    s = bench.schedule(steps)
    assert len(s) == steps and all(0 <= f < bench.FRAMES for f in s)
    if steps <= bench.FRAMES:
        # evenly spaced, distinct, starting at frame 0
        assert s[0] == 0 and len(set(s)) == steps
        gaps = {b - a for a, b in zip(s, s[1:])}
        assert gaps <= {bench.FRAMES // steps, -(-bench.FRAMES // steps)}
        assert bench.FRAMES - s[-1] <= -(-bench.FRAMES // steps)
    else:
        assert s[:bench.FRAMES] == list(range(bench.FRAMES))


def test_schedule_of_twenty_steps():
    assert bench.schedule(20) == list(range(0, 120, 6))


def test_profile_values_present():
    """The committed summaries carry the lines the bench line reads."""
    for name, key in ((bench.HBM_PROFILE, "traffic_per_algorithmic_byte"),
                      (bench.COUNTER_PROFILE, "valu_busy"),
                      (bench.COUNTER_PROFILE, "active_lanes_per_valu"),
                      (bench.COUNTER_PROFILE, "valu_issue_ceiling_gsamples_per_s"),
                      (bench.COUNTER_PROFILE, "l1_lines_per_voxel_load"),
                      (bench.COUNTER_PROFILE, "texture_addresser_busy"),
                      (bench.COUNTER_PROFILE, "l2_hit_rate"),
                      ("composite_8layer_1080p_hbm.txt", "traffic_bytes_per_launch")):
        v = bench.profile_value(name, key)
        assert v is not None and v > 0, (name, key)
    assert bench.profile_value("no_such_file.txt", "x") is None


def test_roofline_limiter_reads_the_counter_summary():
    txt = bench.roofline_limiter()
    assert "?" not in txt and bench.COUNTER_PROFILE in txt
    assert "%.1f" % bench.profile_value(bench.COUNTER_PROFILE, "active_lanes_per_valu") in txt


def test_valu_ceiling_entry():
    e = bench.valu_ceiling_entry(400.0)
    assert e["bound"] == "valu" and 0.5 < e["frac"] < 1.0
    assert abs(e["frac"] * e["ceiling_gsamples_per_s"] - 400.0) < 1e-9


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
