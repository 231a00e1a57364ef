"""Host logic of the batch workload and the autotuner (no GPU)."""
import json
import math
import os

import numpy as np

from rtxplay_amd import autotune
from rtxplay_amd.scene import Frand48, batch_frame, BATCH_PARAM_SEED


def test_batch_frames_follow_the_recipe():
    fr = [batch_frame(i) for i in range(64)]
    assert [f["seed"] for f in fr] == list(range(64))
    for f in fr:
        assert 3 <= f["half"] <= 11 and 2 <= f["level"] <= 6
        assert 320 <= f["w"] <= 1920 and 180 <= f["h"] <= 1080
        assert abs(f["w"] / f["h"] - 16 / 9) < 0.02
        assert 4 <= f["spp"] <= 64 and f["depth"] == 16
        e = f["camera"]["eye"]
        assert abs(math.hypot(e[0], e[2]) - 13.0) < 1e-9 and 1.0 <= e[1] <= 4.0
    # uneven: the spread the workload is about
    px = np.array([f["w"] * f["h"] for f in fr])
    assert px.max() / px.min() > 15 and max(f["spp"] for f in fr) / min(f["spp"] for f in fr) > 8
    assert len({f["level"] for f in fr}) == 5 and len({f["half"] for f in fr}) == 9
    # deterministic, and frame i takes draws 6i..6i+5 of one stream
    assert batch_frame(17) == fr[17]
    r = Frand48(BATCH_PARAM_SEED)
    d = [r() for _ in range(6 * 3)]
    assert fr[2]["half"] == 3 + min(int(d[12] * 9), 8) and fr[2]["level"] == 2 + min(int(d[13] * 5), 4)


def test_coordinate_descent_finds_separable_minimum():
    space = {"a": (1, 2, 3, 4), "b": (10, 20, 30), "c": (5, 6)}
    calls = []

    def cost(cfg):
        calls.append(dict(cfg))
        return (cfg["a"] - 3) ** 2 + abs(cfg["b"] - 20) / 10 + (cfg["c"] == 5) * 0.5

    best, t, hist = autotune.coordinate_descent(cost, {"a": 1, "b": 10, "c": 6}, space, rounds=3)
    assert best == {"a": 3, "b": 20, "c": 6} and t == 0
    # every configuration measured once
    keys = [tuple(sorted(c.items())) for c in calls]
    assert len(keys) == len(set(keys)) == len(hist)


def test_coordinate_descent_keeps_start_on_ties():
    best, _, _ = autotune.coordinate_descent(lambda c: 1.0, {"x": 2}, {"x": (1, 2, 3)})
    assert best == {"x": 2}


def test_tuning_json_roundtrip(tmp_path):
    p = str(tmp_path / "sub" / "tuned.json")
    assert autotune.load(p, "gfx950", "batch64") is None
    tun = dict(tile_w=16, stack_lds=32, blocks_per_cu=4, shade_batch=12, inner_frac=36)
    autotune.save(p, "gfx950", "batch64", tun, 0.5, note="x")
    autotune.save(p, "gfx950", "frame", dict(tun, tile_w=8), 1.0)
    assert autotune.load(p, "gfx950", "batch64") == tun
    assert autotune.load(p, "gfx950", "frame")["tile_w"] == 8
    assert autotune.load(p, "gfx942", "frame") is None
    with open(p) as f:
        assert json.load(f)["gfx950"]["batch64"]["seconds"] == 0.5
