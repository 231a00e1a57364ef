"""Benchmark of the MI355X ray-tracing hot path.

One step = one render of the RTOW final scene with every sphere tessellated
at level 6 (7.95M world-space triangles, seed 0), 1920x1080, 500 spp, depth 16
(the render of optx/camera_r.cu + optx/optics_r.cu under one launch), with the
acceleration structure already built and resident in HBM.

Prints ONE JSON line on stdout (progress goes to stderr).  value = traced rays
per second (hits + misses, counted as the reference's rays-per-pixel AOV) in
Mrays/s; samples/s beside it.  roofline: algorithmic bytes per ray (4-wide
node visits x 112 B + triangle tests x 48 B + shading fetch 80 B + ray in/out
48 B, visits and tests counted by a statistics launch) x rays/s against the
8 TB/s HBM peak.  cpu_baseline: the
oracle's renderer (C, same scene, same semantics) on the host cores over a
bounded sample of rows.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload c3|c2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + samples/s at 1080p 500spp (RMSE-matched); batched speedup; bwd/fwd ratio"
WORKLOADS = {
    # name: (level, w, h, spp, depth)
    "c3": (6, 1920, 1080, 500, 16),
    "c2": (4, 1280, 720, 50, 16),
}
HBM_PEAK_GBS = 8000.0
# bytes per 4-wide node visit (6 plane float4 + the refs float4), per
# triangle test, per shading step (hit triangle 48 + instance record 32) and
# per ray for its own input/output (ray 32 in, hit 16 out)
NODE_BYTES, TRI_BYTES, SHADE_BYTES, IO_BYTES = 112, 48, 80, 48


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(level, w, h, spp, depth, seconds, threads):
    """The oracle renderer on host cores over a bounded sample of rows."""
    import oracle
    from rtxplay_amd.scene import sphere_things
    things = sphere_things(0, 11)
    meshes = {level: oracle.sphere_mesh(1.0, level)}
    tris, inst = oracle.flatten(meshes, things, [level] * len(things))
    log("cpu baseline: oracle BVH over %d triangles" % len(tris))
    sc = oracle.Scene(tris, inst, things)
    cam = oracle.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0)
    order = np.random.default_rng(0).permutation(h // threads)
    rays, rows, t0 = 0, 0, time.perf_counter()
    for band in order:
        y0 = int(band) * threads
        rays += sc.render(cam, w, h, spp, depth, y0=y0, y1=y0 + threads, nthreads=threads)["rays"]
        rows += threads
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%d random rows of %dx%d at %d spp depth %d on the same level-%d scene (%.1f s, %d rays)"
                      % (rows, w, h, spp, depth, level, dt, rays)}


def cpu_reference(seconds):
    """The reference's own CPU tracer, rtow.cxx as built from its sources
    (oracle/_ref/rtow: FP64, analytic spheres, 1280x720, 10 spp, depth 50,
    one thread, as written), for `seconds`: it reports each row it starts
    on stderr, so the rows finished and the time they took are exact; the
    rays of those rows come from the oracle's restatement, which
    reproduces the tracer's image byte for byte.  None when the binary is
    not built."""
    import select
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "rtow")
    if not os.path.exists(exe):
        return None
    import oracle
    w, h = 1280, 720
    p = subprocess.Popen([exe], stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    t0 = time.perf_counter()
    buf, starts = b"", []
    try:
        while time.perf_counter() - t0 < seconds:
            r, _, _ = select.select([p.stderr], [], [], 0.5)
            if not r:
                if p.poll() is not None:
                    break
                continue
            chunk = os.read(p.stderr.fileno(), 4096)
            if not chunk:
                break
            now = time.perf_counter() - t0
            buf += chunk
            parts = buf.split(b"\r")
            buf = parts[-1]
            for tok in parts[:-1]:
                if tok.strip():
                    starts.append(now)
    finally:
        p.kill()
        p.wait()
    done = max(len(starts) - 1, 0)          # rows finished: a later row has started
    if done == 0:
        return None
    dt = starts[done]
    rays = oracle.rtow_rays(done * w, w, h, 10, 50)
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "reference",
            "sample": "rtow.cxx built from the reference sources (g++ -O2): its first %d rows of 1280x720 at "
                      "10 spp, depth 50, FP64 analytic spheres (%.1f s, %d rays)" % (done, dt, rays)}


TUNED_JSON = os.path.join(ROOT, "rtxplay_amd", "tuned.json")
ARCH = "gfx950"


def batched_leg(reps, retune, log_tune):
    """The 64-frame batch: 64 scenes (seeds 0..63) and cameras of uneven size
    rendered by one hipGraph-captured persistent launch with autotuned launch
    parameters, against the same 64 frames as 64 plain launches back to back
    on one stream (the plain path's default parameters).  Returns the record."""
    import rtxplay_amd as rtx
    from rtxplay_amd import autotune
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import BatchFrames, batch_frame

    t0 = time.perf_counter()
    bf = BatchFrames([batch_frame(i) for i in range(64)])
    log("batch: 64 scenes, %d triangles, %d samples, built in %.1f s"
        % (bf.triangles(), bf.samples(), time.perf_counter() - t0))
    stream = Stream()

    def timed(fn):
        fn()
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    def sequential():
        for l in bf.launchers:
            l.ignite(stream=stream, sync=False)

    seq_ms = timed(sequential)
    rays = bf.rays()
    log("batch: 64 plain launches %.1f ms (%d rays)" % (seq_ms, rays))

    ctx = bf.scenes[0]
    default = {k: ctx.tuning()[k] for k in autotune.KEYS}
    b = rtx.Batch(ctx, bf.lps, use_graph=True)
    def_ms = timed(lambda: b.launch(stream, sync=False))
    b.free()
    log("batch: one launch, default tuning %.1f ms" % def_ms)

    tuned = None if retune else autotune.load(TUNED_JSON, ARCH, "batch64")
    source = "stored"
    if tuned is None:
        # on the 64 frames themselves: a quarter-spp proxy ranked the shading
        # batch sizes differently from the full frames
        tuned, tsec, hist = autotune.tune_batch(ctx, bf.lps, log=log_tune)
        autotune.save(TUNED_JSON, ARCH, "batch64", tuned, tsec,
                      note="coordinate descent over %d configurations on the 64 frames" % len(hist))
        source = "autotuned now (%d configurations)" % len(hist)
    ctx.tuning(**tuned)
    b = rtx.Batch(ctx, bf.lps, use_graph=True)
    bat_ms = timed(lambda: b.launch(stream, sync=False))
    b.free()
    log("batch: one launch, tuned %s: %.1f ms" % (tuned, bat_ms))
    # the batch reproduces the plain launches: rays per pixel summed over all frames
    rays_b = bf.rays()
    rec = {"frames": 64, "triangles": bf.triangles(), "samples": bf.samples(), "rays": rays,
           "sequential_ms": round(seq_ms, 2), "batched_default_ms": round(def_ms, 2), "batched_ms": round(bat_ms, 2),
           "speedup": round(seq_ms / bat_ms, 3), "Mrays_per_s_batched": round(rays / (bat_ms * 1e-3) / 1e6, 1),
           "Mrays_per_s_sequential": round(rays / (seq_ms * 1e-3) / 1e6, 1), "tuning": tuned,
           "tuning_source": source, "same_rays": rays_b == rays}
    bf.close()
    return rec


def diff_leg(reps):
    """The differentiable frame: level-4 scene, 640x360, 64 spp, depth 8.
    Forward = one render (rtx_launch); backward = one adjoint path-replay
    launch (rtx_launch_adjoint) for an L2 image loss against a target
    rendered with seeded materials.  Returns the timing record."""
    import ctypes
    import rtxplay_amd as rtx
    from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import RTOWScene

    w, h, spp, depth = 640, 360, 64, 8
    sc = RTOWScene(seed=0, half=11, level=4)
    sc.load()
    handle = sc.build()
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    la = rtx.Launcher(sc, lp)
    types = [t["type"] for t in sc.things_spec]
    m0 = sc.get_materials()
    g = np.random.default_rng(7)
    mt = m0.copy()
    for k, t in enumerate(types):
        if t in (TYPE_DIFFUSE, TYPE_REFLECT):
            mt[k, :3] = np.clip(m0[k, :3] * g.uniform(.6, 1.3, 3), .02, 1.)
        if t == TYPE_REFLECT:
            mt[k, 3] = g.uniform(0, .5)
    sc.set_materials(mt)
    la.ignite()
    T = la.raw_rgb()
    sc.set_materials(m0)
    la.ignite()
    I = la.raw_rgb()
    dldi = rtx.DeviceBuffer.from_array((2.0 * (I - T)).astype(np.float32))
    dpar = rtx.DeviceBuffer(16 * len(types))
    stream = Stream()

    def fwd():
        la.ignite(stream=stream, sync=False)

    def bwd():
        rtx.api.call("rtx_launch_adjoint", sc.ctx, ctypes.byref(lp), dldi.ptr, dpar.ptr, stream.ptr)

    def timed(fn):
        fn()
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    f_ms = timed(fwd)
    b_ms = timed(bwd)
    nparam = sum(3 for t in types if t == TYPE_DIFFUSE) + sum(4 for t in types if t == TYPE_REFLECT)
    log("diff: forward %.2f ms, adjoint %.2f ms, ratio %.2f (%d material parameters)"
        % (f_ms, b_ms, b_ms / f_ms, nparam))
    rec = {"workload": "rtow-final-scene level-4 spheres, 640x360, 64 spp, depth 8, L2 image loss",
           "forward_ms": round(f_ms, 3), "adjoint_ms": round(b_ms, 3), "ratio": round(b_ms / f_ms, 3),
           "parameters": nparam}
    dldi.free()
    dpar.free()
    la.free()
    sc.close()
    return rec


def interactive_leg(frames=120):
    """Repeated launches as the interactive viewer makes them
    (optx/simpleui.cxx:207-245 with optx/simplesm.cxx:978-985): one frame =
    a camera update, Launcher::ignite (synchronous) at 1 spp and pp_sRGB,
    1280x720, depth 16, on the level-6 scene; the camera orbits by 0.5 degree
    per frame.  Reports frames/s, the wall time per frame and the kernel
    time per frame (HIP events), whose difference is the per-launch overhead."""
    import math
    import rtxplay_amd as rtx
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import RTOWScene

    w, h = 1280, 720
    sc = RTOWScene(seed=0, half=11, level=6)
    sc.load()
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, 1, 16, sc.build()
    cam = rtx.Camera()

    def aim(i):
        a = math.atan2(3.0, 13.0) + math.radians(0.5 * i)
        r = math.hypot(13.0, 3.0)
        cam.set((r * math.cos(a), 2.0, r * math.sin(a)), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0)
        lp.camera = cam.c

    la = rtx.Launcher(sc, lp)
    img = rtx.DeviceBuffer(4 * w * h)
    stream = Stream()
    aim(0)
    la.ignite(stream=stream)
    rtx.pp_sRGB(lp.rawRGB, img.ptr, w, h, stream)
    stream.synchronize()
    ev = [(Event(), Event()) for _ in range(frames)]
    t0 = time.perf_counter()
    for i in range(frames):
        aim(i + 1)
        ev[i][0].record(stream)
        la.ignite(stream=stream, sync=False)
        ev[i][1].record(stream)
        rtx.pp_sRGB(lp.rawRGB, img.ptr, w, h, stream)
        stream.synchronize()        # Launcher::ignite is synchronous (optx/launcher.cxx:70)
    wall = (time.perf_counter() - t0) / frames
    kern = sum(e0.elapsed_ms(e1) for e0, e1 in ev) / frames
    rays = int(la.rays_per_pixel().sum(dtype=np.uint64))
    log("interactive: %.2f ms/frame wall, %.2f ms kernel, %.0f frames/s" % (wall * 1e3, kern, 1.0 / wall))
    rec = {"workload": "rtow-final-scene level-6 spheres, 1280x720, 1 spp, depth 16, orbiting camera, "
                       "synchronous launch + sRGB per frame", "frames": frames,
           "frames_per_s": round(1.0 / wall, 1), "ms_per_frame": round(wall * 1e3, 3),
           "kernel_ms_per_frame": round(kern, 3), "overhead_ms_per_frame": round(wall * 1e3 - kern, 3),
           "rays_per_frame": rays}
    img.free()
    la.free()
    sc.close()
    return rec


def two_level_leg(spp=64, steps=3):
    """Flat vs two-level layout (the reference's IAS over one GAS per mesh,
    optx/scene.cxx:38-294) on the level-6 scene and on the reference's
    mixed-level scene (per-sphere levels 9/8/6/3, optx/rtwo.cxx:137-244),
    1920x1080, depth 16, at `spp` per frame: build time, device bytes,
    rays/s, and for the two-level scene the refit (Scene::update) after
    every small sphere moved."""
    import rtxplay_amd as rtx
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import RTOWScene, transform_for

    w, h = 1920, 1080
    rec = {"workload": "rtow-final-scene, 1920x1080, %d spp, depth 16, seed 0" % spp}
    for level in (6, "mixed"):
        for tl in (False, True):
            sc = RTOWScene(seed=0, half=11, level=level, two_level=tl)
            sc.load()
            lp = rtx.LpGeneral()
            lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, 16, sc.build()
            st = sc.stats()
            lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
            la = rtx.Launcher(sc, lp)
            stream = Stream()
            la.ignite(stream=stream)
            e0, e1 = Event(), Event()
            e0.record(stream)
            for _ in range(steps):
                la.ignite(stream=stream, sync=False)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / steps
            rays = int(la.rays_per_pixel().sum(dtype=np.uint64))
            r = {"instanced_triangles": st["num_prims"], "stored_triangles": st["unique_prims"],
                 "build_ms": round(st["build_ms"], 1), "device_bytes": st["device_bytes"],
                 "ms_per_frame": round(ms, 2), "mrays": round(rays / (ms * 1e-3) / 1e6, 1)}
            if tl:
                for k, t in enumerate(sc.things_spec):
                    if t["kind"] == 1:
                        m = dict(t)
                        m["center"] = (m["center"][0], np.float32(m["center"][1] + .05), m["center"][2])
                        sc.set(k, transform_for(m))
                sc.update(sc.handle)
                r["refit_ms"] = round(sc.stats()["build_ms"], 3)
            key = "level%s_%s" % (level, "two_level" if tl else "flat")
            log("%s: %s" % (key, r))
            rec[key] = r
            la.free()
            sc.close()
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stats-spp", type=int, default=16)
    ap.add_argument("--shade-batch", type=int, default=0, help="render tuning (0: the library default)")
    ap.add_argument("--legs", default="frame,batch,diff,interactive,two_level",
                    help="comma list of: frame (the headline render), batch (64 frames), diff (adjoint), "
                         "interactive (repeated 1-spp launches), two_level (flat vs two-level layouts)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-frame batched leg")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable (adjoint) leg")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--retune", action="store_true", help="re-run the batch autotuner (else use the stored tuning)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only the post-process kernel over 64M pixels (768 MiB read, 256 MiB written): "
                         "a plain streaming kernel of known size that calibrates the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (pixels would be sharded "
              "by row bands with the BVH replicated per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import rtxplay_amd as rtx

    level, w, h, spp, depth = WORKLOADS[args.workload]
    if rtx.device_count() < 1:
        sys.exit("bench.py needs a GPU")
    if args.copy_probe:
        n = 64 << 20
        src, dst = rtx.DeviceBuffer(12 * n), rtx.DeviceBuffer(4 * n)
        src.zero()
        for _ in range(3):
            rtx.pp_none(src.ptr, dst.ptr, 8192, n // 8192)
        rtx.synchronize()
        print(json.dumps({"copy_probe": {"pixels": n, "read_bytes": 12 * n, "write_bytes": 4 * n}}))
        return
    legs = set(x for x in args.legs.split(",") if x)
    if args.no_batch:
        legs.discard("batch")
    if args.no_diff:
        legs.discard("diff")
    fr = frame_leg(args, level, w, h, spp, depth) if "frame" in legs else None
    batched = None
    if "batch" in legs:
        batched = batched_leg(args.batch_reps, args.retune, log)
        log("batched speedup %.2fx" % batched["speedup"])
    diff = diff_leg(max(3, args.batch_reps)) if "diff" in legs else None
    inter = interactive_leg() if "interactive" in legs else None
    two = two_level_leg() if "two_level" in legs else None

    cpu = cpu_ref = None
    if not args.no_cpu and fr is not None:
        threads = max(1, min(16, os.cpu_count() or 1))
        cpu = cpu_baseline(level, w, h, spp, depth, args.cpu_seconds, threads)
        log("cpu baseline: %.2f Mrays/s on %d threads" % (cpu["value"], threads))
        cpu_ref = cpu_reference(args.cpu_seconds)
        if cpu_ref:
            log("reference cpu tracer: %.2f Mrays/s on 1 thread" % cpu_ref["value"])

    fr = fr or {}
    out = {
        "metric": METRIC, "value": fr.get("mrays"), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fr.get("ms"), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "rtow-final-scene level-%d spheres (%s triangles), %dx%d, %d spp, depth %d, seed 0"
                               % (level, fr.get("triangles"), w, h, spp, depth),
                   "image": [w, h], "spp": spp, "depth": depth, "triangles": fr.get("triangles"),
                   "parallelism": "single-gpu"},
        "samples_per_s": fr.get("samples_per_s"), "Msamples_per_s": fr.get("Msamples_per_s"),
        "rays_per_frame": fr.get("rays"), "rays_per_sample": fr.get("rays_per_sample"),
        "bvh": fr.get("bvh"), "roofline": fr.get("roofline"),
        "cpu_baseline": cpu, "cpu_reference": cpu_ref,
        "batched_speedup": batched["speedup"] if batched else None, "batched": batched,
        "bwd_fwd_ratio": diff["ratio"] if diff else None, "differentiable": diff,
        "interactive": inter, "two_level": two,
    }
    print(json.dumps(out), flush=True)


def frame_leg(args, level, w, h, spp, depth):
    """The headline frame: build, warmup, K timed renders, then a statistics
    launch at reduced spp for the per-ray node/triangle counts."""
    import rtxplay_amd as rtx
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import RTOWScene

    t0 = time.perf_counter()
    scene = RTOWScene(seed=0, half=11, level=level)
    scene.load()
    handle = scene.build()
    if args.shade_batch:
        scene.tuning(shade_batch=args.shade_batch)
    st = scene.stats()
    log("scene: %d things, %d triangles, %d nodes, depth %d, SAH %.1f, build %.1f ms (host %.1f s)"
        % (len(scene.things_spec), st["num_prims"], st["num_nodes"], st["max_depth"], st["sah_cost"],
           st["build_ms"], time.perf_counter() - t0))

    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    launcher = rtx.Launcher(scene, lp)
    stream = Stream()
    for i in range(args.warmup):
        launcher.ignite(stream=stream, sync=False)
        stream.synchronize()
        log("warmup %d done" % i)
    rtx.synchronize()
    e0, e1 = Event(), Event()
    wall0 = time.perf_counter()
    e0.record(stream)
    for i in range(args.steps):
        launcher.ignite(stream=stream, sync=False)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - wall0
    ms = e0.elapsed_ms(e1) / max(args.steps, 1)
    rpp = launcher.rays_per_pixel()
    rays = int(rpp.sum(dtype=np.uint64))
    log("timed %d steps: %.1f ms/step (wall %.2f s), %d rays/frame" % (args.steps, ms, wall, rays))
    mrays = rays / (ms * 1e-3) / 1e6
    samples = w * h * spp / (ms * 1e-3)

    # bytes per ray from a statistics launch at reduced spp (per-ray averages)
    lp.spp = args.stats_spp
    nodes, tris = launcher.ignite_counted(stream=stream)
    srays = int(launcher.rays_per_pixel().sum(dtype=np.uint64))
    lp.spp = spp
    n_node = float(nodes.sum(dtype=np.uint64)) / srays
    n_tri = float(tris.sum(dtype=np.uint64)) / srays
    b_ray = n_node * NODE_BYTES + n_tri * TRI_BYTES + SHADE_BYTES + IO_BYTES
    achieved = b_ray * rays / (ms * 1e-3) / 1e9
    log("per ray: %.1f node visits, %.1f triangle tests -> %.0f B/ray, %.0f GB/s" % (n_node, n_tri, b_ray, achieved))
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "render_%s_traffic.json" % args.workload)
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
    launcher.free()
    scene.close()
    return {"mrays": round(mrays, 2), "ms": round(ms, 3), "triangles": st["num_prims"],
            "samples_per_s": round(samples, 1), "Msamples_per_s": round(samples / 1e6, 2), "rays": rays,
            "rays_per_sample": round(rays / (w * h * spp), 4),
            "bvh": {"nodes": st["num_nodes"], "leaves": st["num_leaves"], "max_depth": st["max_depth"],
                    "sah_cost": round(st["sah_cost"], 2), "build_ms": round(st["build_ms"], 1)},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                         "bytes_per_ray": round(b_ray, 1), "node_visits_per_ray": round(n_node, 2),
                         "tri_tests_per_ray": round(n_tri, 2)}}


if __name__ == "__main__":
    main()
