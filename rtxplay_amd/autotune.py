"""Autotuner for the render launch parameters.

The render kernels take five launch parameters (rtx_tuning in include/rtx.h):
the pixel tile shape (tile_w x 64/tile_w), the LDS part of the per-lane
traversal stack (stack_lds), the persistent grid (blocks_per_cu), and the two
scheduling thresholds of the traversal/shading loop (shade_batch,
inner_frac).  None of them changes a pixel; all of them change the time.
This module searches them by coordinate descent on measured time: starting
from the current values, it sweeps one parameter at a time over its
candidates with the others fixed, keeps the fastest, and repeats the round
until a round changes nothing (or `rounds` is reached).  Results persist as
JSON keyed by device architecture and workload name.
"""
import json
import os
import time

# candidate values per parameter (max blocks_per_cu is added at run time)
SPACE = {
    "tile_w": (2, 4, 8, 16, 32),
    "stack_lds": (16, 32),
    "shade_batch": (8, 12, 16, 24, 32, 40, 48, 56, 64),
    "inner_frac": (24, 30, 36, 42, 48),
}
KEYS = ("tile_w", "stack_lds", "blocks_per_cu", "shade_batch", "inner_frac")


def coordinate_descent(cost, start, space, rounds=2, log=None):
    """Minimise cost(config) -> seconds over the product of space's
    candidate lists, one coordinate at a time.  Returns (best config, best
    cost, history [(config, cost)]).  Every distinct config is measured once."""
    seen = {}

    def measure(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = float(cost(dict(cfg)))
            if log:
                log("tune %s -> %.4f s" % (" ".join("%s=%d" % kv for kv in sorted(cfg.items())), seen[key]))
        return seen[key]

    best = dict(start)
    best_t = measure(best)
    for _ in range(max(1, rounds)):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best)
                cand[name] = v
                t = measure(cand)
                if t < best_t:
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    history = [(dict(k), t) for k, t in seen.items()]
    return best, best_t, history


def tune_batch(scene, lps, reps=2, rounds=2, log=None):
    """Tunes one batched launch of the frames lps (LpGeneral records whose
    buffers are allocated) on scene's context: each candidate is applied with
    Scene.tuning, a Batch is created (its tile layout depends on tile_w) and
    launched once untimed and `reps` times timed; the minimum counts.  Leaves
    the best tuning set on the context and returns (best, seconds, history)."""
    from .api import Batch
    from .hip import Event, Stream

    stream = Stream()
    cur = scene.tuning()
    space = dict(SPACE)
    space["blocks_per_cu"] = tuple(range(max(1, cur["max_blocks_per_cu"] - 2), cur["max_blocks_per_cu"] + 1))
    start = {k: cur[k] for k in KEYS}

    def cost(cfg):
        scene.tuning(**cfg)
        b = Batch(scene, lps, use_graph=False)
        try:
            b.launch(stream, sync=False)
            stream.synchronize()
            ts = []
            for _ in range(reps):
                e0, e1 = Event(), Event()
                e0.record(stream)
                b.launch(stream, sync=False)
                e1.record(stream)
                e1.synchronize()
                ts.append(e0.elapsed_ms(e1) * 1e-3)
            return min(ts)
        finally:
            b.free()

    t0 = time.perf_counter()
    best, best_t, hist = coordinate_descent(cost, start, space, rounds=rounds, log=log)
    scene.tuning(**best)
    if log:
        log("tuned in %.1f s over %d configurations: %s" % (time.perf_counter() - t0, len(hist), best))
    return best, best_t, hist


def load(path, arch, workload):
    """Stored tuning for (arch, workload) or None."""
    try:
        with open(path) as f:
            db = json.load(f)
    except (OSError, ValueError):
        return None
    ent = db.get(arch, {}).get(workload)
    if not ent:
        return None
    return {k: int(ent["tuning"][k]) for k in KEYS if k in ent["tuning"]}


def save(path, arch, workload, tuning, seconds, note=""):
    """Adds or replaces the entry (arch, workload) of the JSON file at path."""
    try:
        with open(path) as f:
            db = json.load(f)
    except (OSError, ValueError):
        db = {}
    db.setdefault(arch, {})[workload] = {"tuning": {k: int(tuning[k]) for k in KEYS if k in tuning},
                                         "seconds": float(seconds), "note": note}
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
