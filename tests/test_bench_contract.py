"""The bench line's roofline and its evidence, on the CPU: the committed
counter profile belongs to the kernel sources of this tree (its stamp), the
roofline record derived from it is internally consistent (fractions of the
stated peaks, no fraction above 1 at the profile's own rate), a profile of
other sources makes the record stale instead of reusing it, and more than
one GPU is refused with a message."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

PROFILE = os.path.join(ROOT, "profiles", "k_render_level6-1080p_counters.json")


@pytest.fixture(scope="module")
def profile():
    with open(PROFILE) as f:
        return json.load(f)


def test_counter_profile_is_of_these_sources(profile):
    assert profile["stamp"] == bench.kernel_stamp()


def test_roofline_record_from_profile(profile):
    rays = profile["rays_per_launch"]
    ms = profile["kernel_avg_us"] / 1e3
    rec = bench.roofline("level6-1080p", rays, ms, 2600.0)
    assert rec["bound"] == "valu" and not rec["stale"]
    c = profile["counters_per_launch"]
    per_ray = c["SQ_INSTS_VALU"] / rays
    assert rec["valu_inst_per_ray"] == pytest.approx(per_ray, abs=0.01)
    # at the profiled launch's own duration the rate is SQ_INSTS_VALU / time
    assert rec["achieved"] == pytest.approx(c["SQ_INSTS_VALU"] / (ms * 1e-3) / 1e9, rel=1e-3)
    assert rec["frac"] == pytest.approx(rec["achieved"] / rec["peak"], rel=1e-3)
    assert 0.0 < rec["frac"] <= 1.0
    assert rec["peak"] == pytest.approx(256 * 2.4)
    assert 0.0 < rec["lanes_active"] <= 1.0
    hbm = rec["hbm"]
    assert rec["traffic"] == pytest.approx(profile["hbm_bytes_per_launch"])
    assert hbm["frac"] == pytest.approx(rec["traffic"] / (ms * 1e-3) / 1e9 / 8000.0, rel=1e-3)
    assert 0.0 < hbm["frac"] <= 1.0 and 0.0 < hbm["l2_hit"] < 1.0


def test_profile_of_other_sources_is_stale(monkeypatch):
    monkeypatch.setattr(bench, "kernel_stamp", lambda: "0000000000000000")
    rec = bench.roofline("level6-1080p", 1e9, 500.0, 2600.0)
    assert rec["stale"] and rec["frac"] is None and rec["achieved"] is None and rec["traffic"] is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
