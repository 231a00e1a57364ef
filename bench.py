"""Benchmark of the MI355X ray-tracing hot path.

One step = one render of the RTOW final scene as the reference builds it
(RTWO::load on its own rand() stream: 486 spheres), every sphere tessellated
at level 6 (7.96M world-space triangles), 1920x1080, 500 spp, depth 16
(the render of optx/camera_r.cu + optx/optics_r.cu under one launch), with the
acceleration structure already built and resident in HBM.

Prints ONE JSON line on stdout (progress goes to stderr).  value = traced rays
per second (hits + misses, counted as the reference's rays-per-pixel AOV) in
Mrays/s; samples/s beside it.

roofline: the render kernel is bound by VALU issue (counters: most issue
slots busy, half the lanes idle, HBM at a third of its peak), so `bound` is
"valu": achieved = VALU wave-instructions per ray (from the committed counter
profile of the same kernel build, stamped with a hash of the kernel sources)
x rays per second measured here, against 256 CUs x 2.4 GHz x one wave
instruction per CU per clock.  Beside it: the share of lanes active per VALU
instruction, the HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE
counters (scaled by a copy kernel of known size) against the 8 TB/s peak,
and the bytes per ray that the traversal touches (node visits x 112 B +
triangle tests x 48 B + shading 80 B + ray in/out 48 B: bytes served by
L1, L2 and the MALL, most of them never from HBM).

cpu_baseline: the oracle's renderer (C, same scene, same semantics) on the
host cores over a bounded sample of rows; cpu_reference: the reference's
own rtow.cxx built from its sources; cpu_rtow_400x225: the rtow.cxx tracer
on its own analytic scene (400x225, 10 spp, 16 rays per sample) in FP64 as
written and in FP32 on one and on all threads.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload level6-1080p|level4-720p]
    python bench.py --counters      # the counter profile (rocprofv3 passes)
"""
import argparse
import glob
import hashlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s + samples/s at 1080p 500spp (RMSE-matched); batched speedup; bwd/fwd ratio"
WORKLOADS = {
    # name: (level, w, h, spp, depth)
    "level6-1080p": (6, 1920, 1080, 500, 16),
    "level4-720p": (4, 1280, 720, 50, 16),
}
HBM_PEAK_GBS = 8000.0
# VALU issue: one wave instruction per CU per clock, 256 CUs at the 2.4 GHz peak clock
CUS, PEAK_GHZ = 256, 2.4
VALU_PEAK_GINST = CUS * PEAK_GHZ
# bytes per 4-wide node visit (6 plane float4 + the refs float4), per
# triangle test, per shading step (hit triangle 48 + instance record 32) and
# per ray for its own input/output (ray 32 in, hit 16 out)
NODE_BYTES, TRI_BYTES, SHADE_BYTES, IO_BYTES = 112, 48, 80, 48
PROFILES = os.path.join(ROOT, "profiles")
# the sphere list of every leg but the batch: RTWO::load on the reference's
# own rand() stream (optx/rtwo.cxx:137-244), the scene its batch binary renders
SCENE = "glibc"
FRAME_KERNEL = "k_render<256, false, 32, false,"   # the frame kernel (either top-node staging)
COPY_KERNEL = "k_pp<false>"
# counter passes of the headline frame (each within one pass's block limits)
PASSES = [
    ("sq", ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY",
            "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_ANY", "SQ_CYCLES"]),
    ("cache", ["TCC_HIT_sum", "TCC_MISS_sum", "TCP_TCC_READ_REQ_sum", "TCP_TOTAL_CACHE_ACCESSES_sum"]),
    ("fetch", ["FETCH_SIZE"]),
    ("write", ["WRITE_SIZE"]),
]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_stamp():
    """Hash of the sources and build flags the kernels are compiled from: a
    counter profile holds for the build whose stamp it carries."""
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, "rtxplay_amd", "csrc", "*"))) + [
        os.path.join(ROOT, "rtxplay_amd", "build.py")]
    for f in files:
        h.update(os.path.basename(f).encode())
        h.update(open(f, "rb").read())
    return h.hexdigest()[:16]


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


# ---------------------------------------------------------------------------
# CPU legs
# ---------------------------------------------------------------------------
class OracleFrame:
    """The oracle's scene and camera of a workload (built once, used by the
    CPU baseline and the RMSE reference)."""

    def __init__(self, level, w, h):
        import oracle
        from rtxplay_amd.scene import sphere_things_glibc
        t0 = time.perf_counter()
        things = sphere_things_glibc(11)
        meshes = {level: oracle.sphere_mesh(1.0, level)}
        tris, inst = oracle.flatten(meshes, things, [level] * len(things))
        self.scene = oracle.Scene(tris, inst, things)
        self.cam = oracle.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0)
        self.w, self.h, self.level, self.ntris = w, h, level, len(tris)
        log("oracle scene: %d triangles, BVH in %.1f s" % (len(tris), time.perf_counter() - t0))


def cpu_baseline(of, spp, depth, seconds, threads):
    """The oracle renderer on host cores over a bounded sample of rows."""
    w, h = of.w, of.h
    order = np.random.default_rng(0).permutation(h // threads)
    rays, rows, t0 = 0, 0, time.perf_counter()
    for band in order:
        y0 = int(band) * threads
        rays += of.scene.render(of.cam, w, h, spp, depth, y0=y0, y1=y0 + threads, nthreads=threads)["rays"]
        rows += threads
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    rec = {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": "%d random rows of %dx%d at %d spp depth %d on the same level-%d scene (%.1f s, %d rays)"
                     % (rows, w, h, spp, depth, of.level, dt, rays)}
    rec.update(host_info())
    return rec


def cpu_reference(seconds):
    """The reference's own CPU tracer, rtow.cxx as built from its sources
    (oracle/_ref/rtow: FP64, analytic spheres, 1280x720, 10 spp, depth 50,
    one thread, as written), for `seconds`: it reports each row it starts
    on stderr, so the rows finished and the time they took are exact; the
    rays of those rows come from the oracle's restatement, which
    reproduces the tracer's image byte for byte.  None when the binary is
    not built."""
    import select
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
    rec = {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "reference",
           "sample": "rtow.cxx built from the reference sources (g++ -O2): its first %d rows of 1280x720 at "
                     "10 spp, depth 50, FP64 analytic spheres (%.1f s, %d rays)" % (done, dt, rays)}
    rec.update(host_info())
    return rec


def cpu_rtow_small(threads):
    """rtow.cxx's tracer (analytic spheres, its own scene and
    camera) at 400x225, 10 spp, 16 rays per sample (trace depth 15), in
    three forms: FP64 on the reference's glibc stream, one thread (rtow.cxx
    as written); FP32 on per-pixel streams, one thread; FP32 on all host
    threads."""
    import oracle
    w, h, spp, depth = 400, 225, 10, 15
    rec = {"workload": "rtow.cxx scene, analytic spheres, %dx%d, %d spp, at most 16 rays per sample" % (w, h, spp)}
    rec.update(host_info())
    for name, kw in (("fp64_1thread", dict(fp32=False, nthreads=0)), ("fp32_1thread", dict(fp32=True, nthreads=1)),
                     ("fp32_allcores", dict(fp32=True, nthreads=threads))):
        # whole frames until at least a second has passed (the all-core frame takes ~0.1 s)
        t0, reps, rays = time.perf_counter(), 0, 0
        while reps == 0 or time.perf_counter() - t0 < 1.0:
            r = oracle.rtow_render_ex(w, h, spp, depth, **kw)
            reps += 1
            rays += r["rays"]
        dt = time.perf_counter() - t0
        rec[name] = {"threads": max(1, kw["nthreads"]), "frames": reps, "seconds": round(dt, 3),
                     "Mrays_per_s": round(rays / dt / 1e6, 3),
                     "Msamples_per_s": round(reps * w * h * spp / dt / 1e6, 3),
                     "rays_per_sample": round(r["rays"] / (w * h * spp), 4),
                     "image_mean": [round(float(x), 5) for x in r["mean"].reshape(-1, 3).mean(axis=0)]}
        log("rtow 400x225 %s: %.2f Mrays/s (%.2f s)" % (name, rec[name]["Mrays_per_s"], dt))
    return rec


def rmse_leg(of, gpu_raw, spp, depth, threads, rows=2, ref_spp=4096):
    """RMSE of the GPU frame against a 4096-spp CPU reference of the same
    scene on `rows` rows spread over the image (the oracle with per-pixel
    streams independent of the frame's), beside the RMSE of an independent
    CPU render at the frame's spp and the RMSE that the Monte Carlo
    variance alone predicts, sqrt(mean var * (1/spp + 1/4096))."""
    w, h = of.w, of.h
    ys = [int((k + .5) * h / rows) for k in range(rows)]
    t0 = time.perf_counter()
    ref, cpu, var = [], [], []
    for y in ys:
        r = of.scene.render(of.cam, w, h, ref_spp, depth, y0=y, y1=y + 1, nthreads=threads, want_sumsq=True,
                            seed_base=1 << 24)
        ref.append(r["raw"][y].astype(np.float64))
        var.append(np.maximum(r["sumsq"][y].astype(np.float64) / ref_spp - ref[-1] ** 2, 0))
        c = of.scene.render(of.cam, w, h, spp, depth, y0=y, y1=y + 1, nthreads=threads, seed_base=2 << 24)
        cpu.append(c["raw"][y].astype(np.float64))
    ref, cpu, var = np.array(ref), np.array(cpu), np.array(var)
    g = gpu_raw[ys].astype(np.float64)
    rm_g = float(np.sqrt(((g - ref) ** 2).mean()))
    rm_c = float(np.sqrt(((cpu - ref) ** 2).mean()))
    rm_e = float(np.sqrt(var.mean() * (1.0 / spp + 1.0 / ref_spp)))
    log("rmse vs %d spp: gpu %.5f, cpu %.5f, variance predicts %.5f (%.1f s)" % (ref_spp, rm_g, rm_c, rm_e,
                                                                                 time.perf_counter() - t0))
    return {"reference": "oracle render at %d spp with independent streams, rows %s of %dx%d" % (ref_spp, ys, w, h),
            "rmse_gpu": round(rm_g, 6), "rmse_cpu_same_spp": round(rm_c, 6), "rmse_predicted": round(rm_e, 6),
            "gpu_over_predicted": round(rm_g / rm_e, 4), "spp": spp,
            "hit_t_tolerance": "closest hits are bit-exact on decidable rays; hit t within 1e-5 relative on 99.97 % "
                               "of hits, the rest within 2^-15 of the coordinate magnitude along the ray (FP32 "
                               "coordinates near a 13-unit camera; worst 2.3e-5 relative)"}


# ---------------------------------------------------------------------------
# counter profile (rocprofv3 passes over the headline frame)
# ---------------------------------------------------------------------------
def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _counter_means(outdir, kernel):
    """Mean per dispatch of each counter over the dispatches of `kernel`."""
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("no counter_collection.csv under " + outdir)
    per = {}
    for f in files:
        for r in _csv_rows(f):
            if kernel not in r.get("Kernel_Name", ""):
                continue
            d = per.setdefault(r["Counter_Name"], {})
            key = (f, r.get("Dispatch_Id", ""))
            d[key] = d.get(key, 0.0) + float(r["Counter_Value"])
    return {k: sum(v.values()) / len(v) for k, v in per.items()}, {k: len(v) for k, v in per.items()}


def _kernel_stats(outdir):
    files = glob.glob(os.path.join(outdir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise RuntimeError("no kernel_stats.csv under " + outdir)
    return _csv_rows(files[0])


def counters_mode(args):
    """Runs each counter pass and the kernel-trace pass of the headline
    frame under rocprofv3 in its own process (under a kill timeout), plus
    the copy probe's FETCH_SIZE and WRITE_SIZE passes, and writes
    <out>/k_render_<workload>_counters.json and .txt (copied to profiles/
    for the committed record)."""
    out = os.path.abspath(args.counters_out)
    os.makedirs(out, exist_ok=True)
    bench = os.path.abspath(__file__)
    frame = [sys.executable, bench, "--legs", "frame", "--steps", "1", "--warmup", "0", "--no-cpu", "--stats-spp", "1",
             "--workload", args.workload, "--no-rmse"]
    copy = [sys.executable, bench, "--copy-probe"]
    cmds, res, rays = [], {}, None

    def run(name, pre, cmd):
        d = os.path.join(out, "raw_" + name)
        full = ["timeout", "-s", "KILL", "300", "rocprofv3"] + pre + ["-f", "csv", "-d", d, "-o", "run", "--"] + cmd
        shown = " ".join(full[4:]).replace(sys.executable, "python3").replace(bench, "bench.py").replace(out, "OUT")
        cmds.append(shown)
        log("counters: " + shown)
        p = subprocess.run(full, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        if p.returncode != 0:
            raise RuntimeError("rocprofv3 pass %s failed (%d): %s" % (name, p.returncode, p.stderr[-2000:]))
        return d, p.stdout

    for name, counters in PASSES:
        d, so = run(name, ["--pmc"] + counters, frame)
        for line in so.splitlines():
            if line.startswith("{"):
                rays = json.loads(line).get("rays_per_frame")
        m, n = _counter_means(d, FRAME_KERNEL)
        missing = [c for c in counters if c not in m]
        if missing:
            raise RuntimeError("pass %s: no values of %s for a kernel named %s*" % (name, missing, FRAME_KERNEL))
        res.update(m)
    cf = {}
    for name, ctr in (("copy_fetch", "FETCH_SIZE"), ("copy_write", "WRITE_SIZE")):
        d, _ = run(name, ["--pmc", ctr], copy)
        m, n = _counter_means(d, COPY_KERNEL)
        cf[ctr] = m[ctr] * n[ctr]                 # KiB over all the probe's launches
    d, _ = run("trace", ["--kernel-trace", "--stats"], frame[:5] + ["2"] + frame[6:])
    stats = _kernel_stats(d)
    res_kernel = {}
    tf = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    for r in (_csv_rows(tf[0]) if tf else []):
        if FRAME_KERNEL in r.get("Kernel_Name", ""):
            res_kernel = {"lds_bytes_per_block": int(r["LDS_Block_Size"]),
                          "workgroup": int(r["Workgroup_Size_X"]), "grid": int(r["Grid_Size_X"]),
                          "name": r["Kernel_Name"]}
            break
    npx = 64 << 20
    read_factor = 3 * 12 * npx / (cf["FETCH_SIZE"] * 1024.0)     # known bytes / counted bytes
    write_factor = 3 * 4 * npx / (cf["WRITE_SIZE"] * 1024.0)
    hbm = res["FETCH_SIZE"] * 1024.0 * read_factor + res["WRITE_SIZE"] * 1024.0 * write_factor
    fr = [r for r in stats if FRAME_KERNEL in r.get("Name", r.get("KERNEL_NAME", ""))]
    avg_us = float(fr[0].get("AverageNs", fr[0].get("AVERAGE_NS", 0))) / 1e3 if fr else None
    summary = {"workload": args.workload, "kernel": FRAME_KERNEL, "stamp": kernel_stamp(), "rays_per_launch": rays,
               "counters_per_launch": res, "copy_probe_kib": cf, "read_factor": read_factor,
               "write_factor": write_factor, "hbm_bytes_per_launch": hbm, "kernel_avg_us": avg_us,
               "kernel_resources": res_kernel, "commands": cmds}
    jpath = os.path.join(out, "k_render_%s_counters.json" % args.workload)
    json.dump(summary, open(jpath, "w"), indent=1, sort_keys=True)
    lines = ["# counters of the headline frame kernel %s, workload %s (kernel sources stamp %s)"
             % (FRAME_KERNEL, args.workload, summary["stamp"]), "# commands (each its own process; OUT = the output "
             "directory of: python3 bench.py --counters --counters-out OUT):"]
    lines += ["#   " + c for c in cmds]
    lines.append("# rays per launch %s; values are means per dispatch of %s" % (rays, FRAME_KERNEL))
    lines.append("%-32s %16s %12s" % ("counter", "per launch", "per ray"))
    for k in sorted(res):
        lines.append("%-32s %16.4e %12.3f" % (k, res[k], res[k] / rays if rays else float("nan")))
    lines.append("copy probe (k_pp<false>, 64 Mi pixels, 768 MiB read + 256 MiB written per launch, 3 launches): "
                 "FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB -> read factor %.4f, write factor %.4f"
                 % (cf["FETCH_SIZE"], cf["WRITE_SIZE"], read_factor, write_factor))
    lines.append("HBM bytes per launch %.4e (FETCH_SIZE x 1024 x read factor + WRITE_SIZE x 1024 x write factor)" % hbm)
    s = res
    if rays:
        lines.append("VALU wave instructions per ray %.2f; issue fraction at %d CUs x %.1f GHz: see the bench line"
                     % (s["SQ_INSTS_VALU"] / rays, CUS, PEAK_GHZ))
    lines.append("lanes active per VALU instruction: SQ_THREAD_CYCLES_VALU / SQ_INSTS_VALU / 64 = %.3f"
                 % (s["SQ_THREAD_CYCLES_VALU"] / s["SQ_INSTS_VALU"] / 64))
    lines.append("waves waiting: SQ_WAIT_ANY / SQ_WAVE_CYCLES = %.3f; issuing: SQ_ACTIVE_INST_ANY / SQ_WAVE_CYCLES = %.3f"
                 % (s["SQ_WAIT_ANY"] / s["SQ_WAVE_CYCLES"], s["SQ_ACTIVE_INST_ANY"] / s["SQ_WAVE_CYCLES"]))
    lines.append("L2 hit rate %.3f; L1 requests going to L2 %.3f"
                 % (s["TCC_HIT_sum"] / (s["TCC_HIT_sum"] + s["TCC_MISS_sum"]),
                    s["TCP_TCC_READ_REQ_sum"] / s["TCP_TOTAL_CACHE_ACCESSES_sum"]))
    lines.append("kernel trace: %s average %.1f us over the timed frames" % (FRAME_KERNEL, avg_us or float("nan")))
    if res_kernel:
        lines.append("launch: %s: %d bytes of LDS per block of %d lanes, grid of %d lanes"
                     % (res_kernel["name"][:80], res_kernel["lds_bytes_per_block"], res_kernel["workgroup"],
                        res_kernel["grid"]))
    lines.append("")
    lines.append("kernel trace summary (rocprofv3 --kernel-trace --stats, 2 frames):")
    keys = list(stats[0].keys()) if stats else []
    for r in stats[:12]:
        lines.append("  " + "  ".join("%s=%s" % (k, r[k]) for k in keys))
    open(os.path.join(out, "k_render_%s_counters.txt" % args.workload), "w").write("\n".join(lines) + "\n")
    print(json.dumps({"counters": jpath}))


def roofline(workload, rays, ms, b_ray):
    """The roofline record of the frame kernel from the committed counter
    profile (None fields when it is missing or stamped for other sources)."""
    path = os.path.join(PROFILES, "k_render_%s_counters.json" % workload)
    cache = {"bytes_per_ray": round(b_ray, 1), "GB_per_s": round(b_ray * rays / (ms * 1e-3) / 1e9, 1),
             "note": "bytes the traversal and shading touch per ray, served by L1, L2 and the MALL; "
                     "most never come from HBM"}
    rec = {"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINST, "unit": "G VALU wave-instructions/s",
           "frac": None, "traffic": None, "lanes_active": None, "source": os.path.relpath(path, ROOT),
           "stamp": kernel_stamp(), "stale": True, "cache_level": cache, "hbm": None}
    if not os.path.exists(path):
        return rec
    p = json.load(open(path))
    if p.get("stamp") != rec["stamp"] or not p.get("rays_per_launch"):
        return rec
    c = p["counters_per_launch"]
    per_ray = c["SQ_INSTS_VALU"] / p["rays_per_launch"]
    ach = per_ray * rays / (ms * 1e-3) / 1e9
    traffic = p["hbm_bytes_per_launch"] * rays / p["rays_per_launch"]
    rec.update({"achieved": round(ach, 2), "frac": round(ach / VALU_PEAK_GINST, 4), "traffic": traffic,
                "valu_inst_per_ray": round(per_ray, 2),
                "lanes_active": round(c["SQ_THREAD_CYCLES_VALU"] / c["SQ_INSTS_VALU"] / 64, 4), "stale": False,
                "waves_waiting": round(c["SQ_WAIT_ANY"] / c["SQ_WAVE_CYCLES"], 4),
                "hbm": {"achieved": round(traffic / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "read_factor": round(p["read_factor"], 4), "write_factor": round(p["write_factor"], 4),
                        "l2_hit": round(c["TCC_HIT_sum"] / (c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 4)},
                "profile_kernel_avg_ms": round(p["kernel_avg_us"] / 1e3, 3) if p.get("kernel_avg_us") else None})
    return rec


TUNED_JSON = os.path.join(ROOT, "rtxplay_amd", "tuned.json")
ARCH = "gfx950"


def batched_leg(reps, retune, log_tune):
    """The 64-frame batch: 64 scenes (seeds 0..63) and cameras of uneven size
    rendered by one hipGraph-captured persistent launch, against the same 64
    frames as 64 plain launches back to back on one stream (the plain path's
    default parameters).  The batch runs with the library's default tuning
    and with the autotuner's tuning, which is chosen on 64 other frames
    (seeds 64..127, the same generator): the frames it is scored on play no
    part in choosing it.  Returns the record."""
    import rtxplay_amd as rtx
    from rtxplay_amd import autotune
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import BatchFrames, batch_frame

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

    tuned = None if retune else autotune.load(TUNED_JSON, ARCH, "batch64-heldout")
    source = "stored (tuned on the held-out frames, seeds 64..127)"
    if tuned is None:
        t0 = time.perf_counter()
        held = BatchFrames([batch_frame(i) for i in range(64, 128)])
        log("batch: held-out frames (seeds 64..127) built in %.1f s" % (time.perf_counter() - t0))
        tuned, tsec, hist = autotune.tune_batch(held.scenes[0], held.lps, log=log_tune)
        autotune.save(TUNED_JSON, ARCH, "batch64-heldout", tuned, tsec,
                      note="coordinate descent over %d configurations on the held-out frames (seeds 64..127)"
                           % len(hist))
        held.close()
        source = "autotuned now on the held-out frames, seeds 64..127 (%d configurations)" % len(hist)

    t0 = time.perf_counter()
    bf = BatchFrames([batch_frame(i) for i in range(64)])
    log("batch: 64 scenes, %d triangles, %d samples, built in %.1f s"
        % (bf.triangles(), bf.samples(), time.perf_counter() - t0))

    def sequential():
        for l in bf.launchers:
            l.ignite(stream=stream, sync=False)

    seq_ms = timed(sequential)
    rays = bf.rays()
    log("batch: 64 plain launches %.1f ms (%d rays)" % (seq_ms, rays))

    ctx = bf.scenes[0]
    b = rtx.Batch(ctx, bf.lps, use_graph=True)
    def_ms = timed(lambda: b.launch(stream, sync=False))
    b.free()
    log("batch: one launch, default tuning %.1f ms" % def_ms)
    ctx.tuning(**tuned)
    b = rtx.Batch(ctx, bf.lps, use_graph=True)
    bat_ms = timed(lambda: b.launch(stream, sync=False))
    b.free()
    log("batch: one launch, tuning %s (%s): %.1f ms" % (tuned, source, bat_ms))
    # the batch reproduces the plain launches: rays per pixel summed over all frames
    rays_b = bf.rays()
    rec = {"frames": 64, "triangles": bf.triangles(), "samples": bf.samples(), "rays": rays,
           "sequential_ms": round(seq_ms, 2), "batched_default_ms": round(def_ms, 2), "batched_ms": round(bat_ms, 2),
           "speedup": round(seq_ms / bat_ms, 3), "speedup_default_tuning": round(seq_ms / def_ms, 3),
           "Mrays_per_s_batched": round(rays / (bat_ms * 1e-3) / 1e6, 1),
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
    sc = RTOWScene(seed=SCENE, half=11, level=4)
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
    stream.synchronize()
    g = dpar.download(np.float32, (len(types), 4)).astype(np.float64)
    dldi.free()
    dpar.free()

    def loss_of(m):
        sc.set_materials(m)
        la.ignite()
        return float(((la.raw_rgb().astype(np.float64) - T) ** 2).sum())

    # gradients against central differences (h = 1e-3) of the GPU render:
    # albedo is exact (a path's structure does not depend on it); fuzz moves
    # directions, and the render's differences then also hold the jumps of
    # paths that cross a facet edge, a silhouette or the metal's absorption
    # boundary, which the adjoint's interior derivative leaves out
    hfd = 1e-3
    for name, col, pool in (("albedo", None, [k for k, t in enumerate(types) if t in (TYPE_DIFFUSE, TYPE_REFLECT)]),
                            ("fuzz", 3, [k for k, t in enumerate(types) if t == TYPE_REFLECT and m0[k, 3] >= 2 * hfd])):
        cand = sorted(((abs(g[k, c]), k, c) for k in pool for c in ((0, 1, 2) if col is None else (col,))),
                      reverse=True)[:8]
        rels = []
        for _, k, c in cand:
            mp, mm = m0.copy(), m0.copy()
            mp[k, c] = np.float32(m0[k, c] + hfd)
            mm[k, c] = np.float32(m0[k, c] - hfd)
            fd = (loss_of(mp) - loss_of(mm)) / (float(mp[k, c]) - float(mm[k, c]))
            rels.append(abs(g[k, c] - fd) / abs(fd))
        rec["%s_vs_render_fd" % name] = {"h": hfd, "params": len(rels), "median_rel_err": float(np.median(rels)),
                                         "max_rel_err": float(max(rels))}
        log("diff: %s adjoint vs render differences: median rel err %.2e, max %.2e"
            % (name, np.median(rels), max(rels)))
    sc.set_materials(m0)
    rec["recovery"] = recover_leg(sc, la, lp, types, m0, mt, T)
    rec["recovery_fuzz_spsa"] = recover_leg(sc, la, lp, types, m0, mt, T, fuzz_spsa=0.02)
    la.free()
    sc.close()
    return rec


def recover_leg(sc, la, lp, types, m0, truth, T, iters=150, lr=0.02, lr_end=0.001, fuzz_spsa=0.0, seed=11):
    """The recovery run of the differentiable frame without torch: Adam on the L2 image loss
    from albedo 0.5 and fuzz 0.25, gradients from rtx_launch_adjoint; the
    mean absolute error of albedo and of fuzz over the parameters the image
    constrains (nonzero gradient at the start), before and after.  With
    fuzz_spsa = c > 0 the fuzz gradient comes instead from simultaneous
    perturbation of all fuzz values by +-c (two more renders per step), a
    difference of the real render that holds the boundary terms."""
    from rtxplay_amd import TYPE_DIFFUSE, TYPE_REFLECT
    mask = np.zeros_like(m0)
    for k, t in enumerate(types):
        if t in (TYPE_DIFFUSE, TYPE_REFLECT):
            mask[k, :3] = 1
        if t == TYPE_REFLECT:
            mask[k, 3] = 1
    q = np.where(mask > 0, np.where(np.arange(4)[None, :] < 3, .5, .25), m0).astype(np.float64)
    m, v = np.zeros_like(q), np.zeros_like(q)
    rng = np.random.default_rng(seed)

    def lossq(p):
        sc.set_materials(p.astype(np.float32))
        la.ignite()
        return float(((la.raw_rgb().astype(np.float64) - T) ** 2).sum())

    t0 = time.perf_counter()
    seen, loss0, loss = None, None, None
    for it in range(iters + 1):
        sc.set_materials(q.astype(np.float32))
        la.ignite()
        img = la.raw_rgb().astype(np.float64)
        loss = float(((img - T) ** 2).sum())
        g = la.adjoint((2.0 * (img - T)).astype(np.float32)).astype(np.float64) * mask
        if it == 0:
            seen = (np.abs(g) > 0) & (mask > 0)
        if fuzz_spsa > 0:
            fz = mask[:, 3] > 0
            delta = np.where(rng.random(len(q)) < .5, -1.0, 1.0) * fz
            lo = q.copy()
            hi = q.copy()
            hi[:, 3] = q[:, 3] + fuzz_spsa * delta
            lo[:, 3] = q[:, 3] - fuzz_spsa * delta
            lh = lossq(hi)
            ll = lossq(lo)
            g[:, 3] = np.where(fz, (lh - ll) / (2 * fuzz_spsa * np.where(fz, delta, 1.0)), 0.0)
        if it == 0:
            loss0 = loss
            e0 = np.abs(q - truth)
        if it == iters:
            break
        a = lr * (lr_end / lr) ** (it / max(iters - 1, 1))
        m = .9 * m + .1 * g
        v = .999 * v + .001 * g * g
        q = np.clip(q - a * (m / (1 - .9 ** (it + 1))) / (np.sqrt(v / (1 - .999 ** (it + 1))) + 1e-8), 0.0, 1.0)
    e1 = np.abs(q - truth)
    alb, fz = seen.copy(), seen.copy()
    alb[:, 3] = False
    fz[:, :3] = False
    rec = {"iters": iters, "seconds": round(time.perf_counter() - t0, 1), "loss_start": round(loss0, 3),
           "loss_end": round(loss, 3), "albedo_params": int(alb.sum()), "fuzz_params": int(fz.sum()),
           "albedo_abs_err_start": round(float(e0[alb].mean()), 4), "albedo_abs_err_end": round(float(e1[alb].mean()), 4),
           "fuzz_abs_err_start": round(float(e0[fz].mean()), 4), "fuzz_abs_err_end": round(float(e1[fz].mean()), 4)}
    log("diff: recovery %s" % rec)
    sc.set_materials(m0)
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
    sc = RTOWScene(seed=SCENE, half=11, level=6)
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
    rec = {"workload": "rtow-final-scene (the reference's rand() stream), 1920x1080, %d spp, depth 16" % spp}
    for level in (6, "mixed"):
        for tl in (False, True):
            sc = RTOWScene(seed=SCENE, half=11, level=level, two_level=tl)
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


def frame_leg(args, level, w, h, spp, depth):
    """The headline frame: build, warmup, K timed renders, then a statistics
    launch at reduced spp for the per-ray node/triangle counts.  Returns the
    record and the last frame's rawRGB."""
    import rtxplay_amd as rtx
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import RTOWScene

    t0 = time.perf_counter()
    scene = RTOWScene(seed=SCENE, half=11, level=level)
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
    raw = launcher.raw_rgb().copy()
    rays = int(rpp.sum(dtype=np.uint64))
    log("timed %d steps: %.1f ms/step (wall %.2f s), %d rays/frame" % (args.steps, ms, wall, rays))
    mrays = rays / (ms * 1e-3) / 1e6
    samples = w * h * spp / (ms * 1e-3)

    # node visits and triangle tests per ray from a statistics launch at reduced spp
    lp.spp = args.stats_spp
    nodes, tris = launcher.ignite_counted(stream=stream)
    srays = int(launcher.rays_per_pixel().sum(dtype=np.uint64))
    lp.spp = spp
    n_node = float(nodes.sum(dtype=np.uint64)) / srays
    n_tri = float(tris.sum(dtype=np.uint64)) / srays
    b_ray = n_node * NODE_BYTES + n_tri * TRI_BYTES + SHADE_BYTES + IO_BYTES
    rf = roofline(args.workload, rays, ms, b_ray)
    rf.update({"node_visits_per_ray": round(n_node, 2), "tri_tests_per_ray": round(n_tri, 2)})
    log("per ray: %.1f node visits, %.1f triangle tests; roofline %s" % (n_node, n_tri, {k: rf[k] for k in (
        "bound", "achieved", "frac", "lanes_active", "stale")}))
    launcher.free()
    scene.close()
    return {"mrays": round(mrays, 2), "ms": round(ms, 3), "triangles": st["num_prims"],
            "samples_per_s": round(samples, 1), "Msamples_per_s": round(samples / 1e6, 2), "rays": rays,
            "rays_per_sample": round(rays / (w * h * spp), 4),
            "bvh": {"nodes": st["num_nodes"], "leaves": st["num_leaves"], "max_depth": st["max_depth"],
                    "sah_cost": round(st["sah_cost"], 2), "build_ms": round(st["build_ms"], 1),
                    "wide_nodes": st["wide_nodes"], "wide_depth": st["wide_depth"],
                    "device_bytes": st["device_bytes"],
                    "note": "nodes, leaves, depth and SAH cost of the binary SAH tree; wide_*: the 4-wide "
                            "tree the kernels traverse (its SAH-optimal grouping)"},
            "roofline": rf}, raw


def level4_leg(steps=5):
    """The level-4 scene (497,664 triangles), 1280x720, 50 spp,
    depth 16: build, one warmup, `steps` timed launches."""
    import rtxplay_amd as rtx
    from rtxplay_amd.hip import Event, Stream
    from rtxplay_amd.scene import RTOWScene
    level, w, h, spp, depth = WORKLOADS["level4-720p"]
    sc = RTOWScene(seed=SCENE, half=11, level=level)
    sc.load()
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, sc.build()
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
    la.free()
    sc.close()
    rec = {"workload": "rtow-final-scene level-4 spheres (%d triangles, the reference's rand() stream), %dx%d, "
                       "%d spp, depth %d"
                       % (st["num_prims"], w, h, spp, depth), "ms_per_frame": round(ms, 3),
           "Mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 1),
           "Msamples_per_s": round(w * h * spp / (ms * 1e-3) / 1e6, 1), "rays_per_frame": rays,
           "build_ms": round(st["build_ms"], 1)}
    log("level4: %s" % rec)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="level6-1080p", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rmse", action="store_true", help="skip the RMSE against the 4096-spp CPU reference")
    ap.add_argument("--stats-spp", type=int, default=16)
    ap.add_argument("--shade-batch", type=int, default=0, help="render tuning (0: the library default)")
    ap.add_argument("--legs", default="frame,level4,batch,diff,interactive,two_level",
                    help="comma list of: frame (the headline render), level4 (the level-4 scene at 720p), "
                         "batch (64 frames), "
                         "diff (adjoint), interactive (repeated 1-spp launches), two_level (flat vs two-level "
                         "layouts)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-frame batched leg")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable (adjoint) leg")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--retune", action="store_true",
                    help="re-run the batch autotuner on the held-out frames (seeds 64..127) and store it")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only the post-process kernel over 64M pixels (768 MiB read, 256 MiB written): "
                         "a plain streaming kernel of known size that calibrates the HBM byte counters")
    ap.add_argument("--counters", action="store_true",
                    help="run the rocprofv3 counter passes of the headline frame and write their summary")
    ap.add_argument("--counters-out", default=os.path.join(ROOT, "counters_out"),
                    help="directory for the raw counter passes and the summary (default: counters_out/)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (pixels would be sharded "
              "by row bands with the BVH replicated per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.counters:
        counters_mode(args)
        return

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
    fr, raw = frame_leg(args, level, w, h, spp, depth) if "frame" in legs else (None, None)
    l4 = level4_leg() if "level4" in legs and args.workload == "level6-1080p" else None
    batched = None
    if "batch" in legs:
        batched = batched_leg(args.batch_reps, args.retune, log)
        log("batched speedup %.2fx" % batched["speedup"])
    diff = diff_leg(max(3, args.batch_reps)) if "diff" in legs else None
    inter = interactive_leg() if "interactive" in legs else None
    two = two_level_leg() if "two_level" in legs else None

    cpu = cpu_ref = rtow_small = rmse = None
    threads = max(1, min(16, os.cpu_count() or 1))
    if fr is not None and not (args.no_cpu and args.no_rmse):
        of = OracleFrame(level, w, h)
        if not args.no_rmse:
            rmse = rmse_leg(of, raw, spp, depth, threads)
        if not args.no_cpu:
            cpu = cpu_baseline(of, spp, depth, args.cpu_seconds, threads)
            log("cpu baseline: %.2f Mrays/s on %d threads" % (cpu["value"], threads))
        del of
    if fr is not None and not args.no_cpu:
        cpu_ref = cpu_reference(args.cpu_seconds)
        if cpu_ref:
            log("reference cpu tracer: %.2f Mrays/s on 1 thread" % cpu_ref["value"])
        rtow_small = cpu_rtow_small(threads)

    fr = fr or {}
    out = {
        "metric": METRIC, "value": fr.get("mrays"), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fr.get("ms"), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "rtow-final-scene level-%d spheres (%s triangles, the reference's rand() stream), "
                               "%dx%d, %d spp, depth %d"
                               % (level, fr.get("triangles"), w, h, spp, depth),
                   "image": [w, h], "spp": spp, "depth": depth, "triangles": fr.get("triangles"),
                   "parallelism": "single-gpu"},
        "samples_per_s": fr.get("samples_per_s"), "Msamples_per_s": fr.get("Msamples_per_s"),
        "rays_per_frame": fr.get("rays"), "rays_per_sample": fr.get("rays_per_sample"),
        "bvh": fr.get("bvh"), "roofline": fr.get("roofline"), "rmse": rmse,
        "cpu_baseline": cpu, "cpu_reference": cpu_ref, "cpu_rtow_400x225": rtow_small, "level4_720p": l4,
        "batched_speedup": batched["speedup"] if batched else None, "batched": batched,
        "bwd_fwd_ratio": diff["ratio"] if diff else None, "differentiable": diff,
        "interactive": inter, "two_level": two,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
