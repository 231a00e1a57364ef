"""GPU parity through the C-ABI: device RNG, build, closest-hit query,
render, picker and post-process against the oracle."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN
from test_oracle import t_within_tolerance

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu]", *a, flush=True)


def _scene(rtx, level, seed=0, half=11):
    from rtxplay_amd.scene import RTOWScene
    s = RTOWScene(seed=seed, half=half, level=level)
    s.load()
    say("build level", level, "half", half, "seed", seed)
    s.build()
    return s


def _oracle_scene(orc, s):
    lv = sorted(set(s.levels))
    meshes = {k: orc.sphere_mesh(1.0, k) for k in lv}
    tris, inst = orc.flatten(meshes, s.things_spec, s.levels)
    return tris, inst


@pytest.fixture(scope="module")
def small(rtx, orc):
    s = _scene(rtx, 2, seed=0, half=3)
    tris, inst = _oracle_scene(orc, s)
    yield s, tris, inst
    s.close()


@pytest.fixture(scope="module")
def level4(rtx, orc):
    s = _scene(rtx, 4)
    tris, inst = _oracle_scene(orc, s)
    yield s, tris, inst
    s.close()


def _rays(n, seed, lo=(-6, .1, -6), hi=(6, 3, 6)):
    g = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = g.uniform(lo, hi, (n, 3))
    rays[:, 3] = 1e-3
    rays[:, 4:7] = g.normal(size=(n, 3))
    rays[:, 7] = 1e16
    return rays


# ---- RNG ---------------------------------------------------------------------
def test_device_frand48_kat(rtx):
    k = np.load(os.path.join(GOLDEN, "frand48_kat.npz"))
    say("rng draws", k["draws"].shape)
    got = rtx.rng_draws(k["seeds"], k["draws"].shape[1])
    assert np.array_equal(got.view(np.uint32), k["draws"].view(np.uint32))


# ---- build -------------------------------------------------------------------
def test_build_flatten_bit_exact(small, rtx):
    s, tris, inst = small
    st = s.stats()
    say("stats", st)
    g_tris, g_inst = s.triangles()
    assert st["num_prims"] == len(tris)
    assert np.array_equal(g_tris.reshape(-1, 9).view(np.uint32), tris.view(np.uint32))
    assert np.array_equal(g_inst, inst)
    assert 1 <= st["max_leaf"] <= 8 and st["max_depth"] <= 64 and st["num_leaves"] >= len(tris) / 8
    assert st["num_nodes"] == st["num_leaves"] - 1
    assert math.isfinite(st["sah_cost"]) and st["sah_cost"] > 0


def test_build_level4_stats(level4):
    s, tris, _ = level4
    st = s.stats()
    say("level4 stats", st)
    assert st["num_prims"] == len(tris) == 1024 * len(s.things_spec)
    assert st["max_depth"] <= 64 and st["num_nodes"] == st["num_leaves"] - 1


def test_build_tree_is_unchanged(level4):
    """The builder's tree for the level-4 scene, pinned: the block-per-task,
    wave-per-task, thread-per-task (tasks of at most 16 primitives on levels
    of 8192 tasks or more) and chunked (wide) level paths must all give this tree
    (binning is sums, minima and maxima, and the partition is stable), so a
    change of any of them that alters the tree fails here.  The values are
    the builder's own output, recorded when the paths were introduced."""
    s, _, _ = level4
    st = s.stats()
    assert (st["num_nodes"], st["num_leaves"], st["max_depth"]) == (291545, 291546, 27)
    assert abs(st["sah_cost"] - 27.92353057861328) < 1e-9


def test_build_is_deterministic(rtx, small):
    s, tris, _ = small
    a = s.trace(_rays(4096, 11))
    s2 = _scene(rtx, 2, seed=0, half=3)
    b = s2.trace(_rays(4096, 11))
    assert s2.stats()["num_nodes"] == s.stats()["num_nodes"]
    assert np.array_equal(a, b)
    s2.close()


def test_empty_and_single_triangle_scenes(rtx):
    s = rtx.Scene()
    say("empty scene")
    h = s.build()
    hits = s.trace(_rays(100, 3))
    assert (hits["instance_id"] == 0xFFFFFFFF).all() and np.isinf(hits["t"]).all()
    m = s.add((np.array([[0, 0, -1], [1, 0, -1], [0, 1, -1]], np.float32), np.array([[0, 1, 2]], np.uint32)))
    s.add(rtx.Thing.diffuse((.5, .5, .5)), m)
    say("single triangle scene")
    s.build()
    ray = np.array([[.2, .3, 0, 1e-3, 0, 0, -1, 1e16]], np.float32)
    hit = s.trace(ray)[0]
    assert hit["instance_id"] == 0 and hit["primitive_index"] == 0 and abs(hit["t"] - 1) < 1e-6
    assert abs(hit["u"] - .2) < 1e-6 and abs(hit["v"] - .3) < 1e-6
    s.close()


def test_set_get_update(rtx):
    s = rtx.Scene()
    v, i = rtx.Sphere(1.0, 2).mesh()
    m = s.add((v, i))
    t = s.add(rtx.Thing.diffuse((.5, .5, .5)), m)
    assert s.get(t).tolist() == [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    assert s.set(t + 5, np.eye(3, 4)) is False and s.get(t + 5) is None
    h = s.build()
    say("update after move")
    ray = np.array([[0, 0, 5, 1e-3, 0, 0, -1, 1e16]], np.float32)
    assert abs(s.trace(ray)[0]["t"] - 4) < 1e-2
    assert s.set(t, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -10])
    h2 = s.update(h)
    assert abs(s.trace(ray, handle=h2)[0]["t"] - 14) < 1e-2
    s.close()


def test_bad_mesh_rejected(rtx):
    s = rtx.Scene()
    with pytest.raises(rtx.RtxError) as e:
        s.add((np.zeros((3, 3), np.float32), np.array([[0, 1, 3]], np.uint32)))
    assert e.value.status == 1
    s.close()


# ---- closest-hit query --------------------------------------------------------
def _check_hits(s, tris, inst, rays, ref64, want_ratio=0.95):
    hits = s.trace(rays)
    ok = __import__("oracle").decidable(ref64, rays, tris)
    say("decidable fraction %.5f of %d rays" % (ok.mean(), len(rays)))
    assert ok.mean() > want_ratio
    gp = np.where(hits["instance_id"] == 0xFFFFFFFF, 0xFFFFFFFF, hits["prim"]).astype(np.uint32)
    bad = ok & (gp != ref64["prim"])
    assert not bad.any(), (np.where(bad)[0][:10], gp[bad][:10], ref64["prim"][bad][:10])
    hit = ok & np.isfinite(ref64["t"])
    assert t_within_tolerance(hits["t"], ref64["t"], rays, hit)
    # instance / primitive index agree with the flattened order
    h = np.isfinite(hits["t"])
    assert np.array_equal(hits["instance_id"][h], inst[hits["prim"][h]])
    return hits, ok


def test_trace_random_rays_vs_fp64_brute_force(small, orc):
    s, tris, inst = small
    rays = _rays(20000, 5)
    say("brute force", len(tris), "tris")
    ref = orc.closest_f64_brute(tris, rays)
    _check_hits(s, tris, inst, rays, ref)


def test_trace_matches_fp32_oracle_exactly(small, orc):
    """Same watertight arithmetic and tie rule: identical hits for every ray."""
    s, tris, inst = small
    rays = _rays(20000, 6)
    ref = orc.BVH(tris).closest_f32(rays)
    hits = s.trace(rays)
    gp = np.where(hits["instance_id"] == 0xFFFFFFFF, 0xFFFFFFFF, hits["prim"]).astype(np.uint32)
    same = gp == ref["prim"]
    say("fp32 agreement", same.mean())
    assert same.mean() > 0.9999
    h = same & (gp != 0xFFFFFFFF)
    assert np.array_equal(hits["t"][h], ref["t"][h])


def test_trace_camera_and_bounce_rays_level4(level4, orc):
    s, tris, inst = level4
    cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 16 / 9, 0.1, 10.0)
    prim = orc.primary_rays(cam, 320, 180)
    bvh = orc.BVH(tris)
    ref = bvh.closest_f64(prim)
    hits, ok = _check_hits(s, tris, inst, prim, ref)
    # first-bounce rays: from the FP64 hit points in mirror directions
    h = np.isfinite(ref["t"])
    o = prim[h, 0:3].astype(np.float64) + ref["t"][h, None] * prim[h, 4:7]
    T = tris[ref["prim"][h]].reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    d = prim[h, 4:7].astype(np.float64)
    d = d / np.linalg.norm(d, axis=1)[:, None]
    r = d - 2 * (d * n).sum(axis=1)[:, None] * n
    b = np.zeros((h.sum(), 8), np.float32)
    b[:, 0:3] = o; b[:, 3] = 1e-3; b[:, 4:7] = r; b[:, 7] = 1e16
    ref_b = bvh.closest_f64(b)
    _check_hits(s, tris, inst, b, ref_b, want_ratio=0.9)


# ---- render ------------------------------------------------------------------
def _launch(rtx, s, w, h, spp, depth, cam=None):
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth = w, h, spp, depth
    lp.is_handle = s.handle
    c = cam or rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0)
    lp.camera = c.c
    L = rtx.Launcher(s, lp)
    say("render %dx%d spp %d depth %d" % (w, h, spp, depth))
    L.ignite()
    return L, lp


def _oracle_render(orc, tris, inst, s, w, h, spp, depth, sumsq=False):
    sc = orc.Scene(tris, inst, s.things_spec)
    cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0)
    return sc.render(cam, w, h, spp, depth, nthreads=8, want_sumsq=sumsq)


def test_render_matches_oracle(rtx, orc, level4):
    s, tris, inst = level4
    w, h, spp, depth = 96, 54, 16, 16
    L, lp = _launch(rtx, s, w, h, spp, depth)
    raw, rpp = L.raw_rgb(), L.rays_per_pixel()
    nrm, alb = L.normals(), L.albedos()
    ref = _oracle_render(orc, tris, inst, s, w, h, spp, depth, sumsq=True)
    # statistical criterion: per-channel image mean within 3 sigma of the MC error
    var = np.maximum(ref["sumsq"] / spp - (ref["raw"]) ** 2, 0) / spp
    sigma = np.sqrt(var.reshape(-1, 3).sum(axis=0)) / (w * h)
    dm = np.abs(raw.reshape(-1, 3).mean(axis=0) - ref["raw"].reshape(-1, 3).mean(axis=0))
    say("channel mean diff", dm, "3 sigma", 3 * sigma)
    assert (dm <= 3 * sigma * np.sqrt(2)).all()
    # same RNG streams and the same roundings (both sides round every
    # operation as written): the pixels are bit-identical
    close = np.abs(raw - ref["raw"]).max(axis=2) < 1e-4
    exact = (raw.view(np.uint32) == ref["raw"].view(np.uint32)).all(axis=2)
    say("pixels within 1e-4:", close.mean(), " bit-identical:", exact.mean(), " rpp equal:", (rpp == ref["rpp"]).mean())
    assert exact.mean() >= 0.999 and close.mean() > 0.999 and (rpp == ref["rpp"]).mean() >= 0.999
    assert abs(int(rpp.sum()) - int(ref["rpp"].sum())) / ref["rpp"].sum() < 0.01
    assert (nrm.view(np.uint32) == ref["normals"].view(np.uint32)).all(axis=2).mean() >= 0.999
    assert (alb.view(np.uint32) == ref["albedos"].view(np.uint32)).all(axis=2).mean() >= 0.999
    assert raw.min() >= 0 and raw.max() <= 1
    L.free()


def test_render_matches_oracle_iterative(rtx, orc, level4):
    """The iterative programs' semantics (rtx_set_semantics, optx/camera_i.cu
    + optx/optics_i.cu) at depth 4, where many samples reach the depth
    limit: the frame equals the oracle's iterative render bit for bit, and
    differs from the recursive frame where a sample ends at the limit."""
    s, tris, inst = level4
    w, h, spp, depth = 96, 54, 8, 4
    L, lp = _launch(rtx, s, w, h, spp, depth)
    rec = L.raw_rgb().copy()
    s.semantics(s.ITERATIVE)
    try:
        L.ignite()
        raw, rpp, nrm, alb = L.raw_rgb(), L.rays_per_pixel(), L.normals(), L.albedos()
    finally:
        s.semantics(s.RECURSIVE)
    sc = orc.Scene(tris, inst, s.things_spec, iterative=True)
    ref = sc.render(orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0), w, h, spp, depth, nthreads=8)
    exact = (raw.view(np.uint32) == ref["raw"].view(np.uint32)).all(axis=2)
    say("iterative: bit-identical", exact.mean(), "pixels differing from recursive", (raw != rec).any(axis=2).mean())
    assert exact.mean() >= 0.999 and (rpp == ref["rpp"]).mean() >= 0.999
    assert (nrm.view(np.uint32) == ref["normals"].view(np.uint32)).all(axis=2).mean() >= 0.999
    assert (alb.view(np.uint32) == ref["albedos"].view(np.uint32)).all(axis=2).mean() >= 0.999
    assert (raw != rec).any(axis=2).mean() > 0.05
    L.free()


def test_render_depth_one_and_limits(rtx, level4):
    s, _, _ = level4
    L, lp = _launch(rtx, s, 64, 36, 4, 1)
    rpp = L.rays_per_pixel()
    assert (rpp == 4).all()                      # depth 1: only the primary ray per sample
    raw = L.raw_rgb()
    # every sample is either black (hit at the depth limit) or sky: blue >= red
    assert (raw[..., 2] >= raw[..., 0] - 1e-6).all()
    lp.spp = 0
    with pytest.raises(rtx.RtxError):
        L.ignite()
    lp.spp, lp.depth = 4, 0
    with pytest.raises(rtx.RtxError):
        L.ignite()
    lp.depth = 16
    lp.is_handle = 12345
    with pytest.raises(rtx.RtxError) as e:
        L.ignite()
    assert e.value.status == 2
    L.free()


def test_render_odd_sizes_and_repeat(rtx, level4):
    s, _, _ = level4
    L, lp = _launch(rtx, s, 37, 23, 3, 16)
    a = L.raw_rgb().copy()
    L.ignite()
    assert np.array_equal(a, L.raw_rgb())         # deterministic: streams re-seeded per launch
    assert (L.rays_per_pixel() >= 3).all()
    L.free()


@pytest.mark.parametrize("tile_w", [8, 4, 32])
def test_longest_first_order_renders_every_pixel_once(rtx, level4, tile_w):
    """Frames of 16 or more spp and at least two pixels per lane of the
    persistent grid (1024 blocks of 256 lanes on 256 CUs) take their tiles
    in descending order of a one-sample probe's rays; the counted launch
    keeps the slot order.  On an odd frame size (partial edge tiles) the
    ordered launch writes every pixel (each buffer is pre-filled with a
    sentinel) and gives the counted launch's pixels bit for bit."""
    s, _, _ = level4
    old = s.tuning()
    try:
        s.tuning(tile_w=tile_w)
        w, h, spp = 1283, 719, 16
        lp = rtx.LpGeneral()
        lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, 16, s.handle
        lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
        L = rtx.Launcher(s, lp)
        L.ignite_counted()
        ref = (L.raw_rgb(), L.rays_per_pixel(), L.normals(), L.albedos())
        for k in ("rawRGB", "normals", "albedos"):
            L._bufs[k].upload(np.full((h, w, 3), np.nan, np.float32))
        L._bufs["rpp"].upload(np.full((h, w), 0xFFFFFFFF, np.uint32))
        L.ignite()
        got = (L.raw_rgb(), L.rays_per_pixel(), L.normals(), L.albedos())
        say("tile_w", tile_w, "ordered launch vs slot order: rays", int(got[1].sum()), int(ref[1].sum()))
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        L.free()
    finally:
        s.tuning(**{k: old[k] for k in ("tile_w", "stack_lds", "blocks_per_cu", "shade_batch", "inner_frac")})


def test_picker(rtx, orc, level4):
    s, tris, inst = level4
    w, h = 320, 180
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth = w, h, 1, 16
    lp.is_handle = s.handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    L = rtx.Launcher(s, lp)
    cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0)
    prim = orc.primary_rays(cam, w, h)
    ref = orc.BVH(tris).closest_f32(prim)
    for (x, y) in [(160, 90), (10, 170), (300, 20), (5, 5)]:
        lp.picker, lp.pick_x, lp.pick_y = True, x, y
        say("pick", x, y)
        L.ignite(once=True)
        p = ref["prim"][w * y + x]
        want = 0xFFFFFFFF if p == 0xFFFFFFFF else int(inst[p])
        assert L.pick_id() == want
    L.free()


def test_postprocess_matches_oracle(rtx, orc):
    g = np.random.default_rng(9)
    w, h = 257, 33
    raw = g.uniform(0, 1, (h, w, 3)).astype(np.float32)
    raw[0, :5] = [[0, 0, 0], [1, 1, 1], [.0031308, .0031307, .5], [1e-9, .25, .75], [.999, .001, .1]]
    src = rtx.DeviceBuffer.from_array(raw)
    dst = rtx.DeviceBuffer(4 * w * h)
    rtx.pp_sRGB(src.ptr, dst.ptr, w, h)
    rtx.synchronize()
    got = dst.download(np.uint8, (h, w, 4))
    want = orc.pp_srgb(raw)
    diff = np.abs(got.astype(int) - want.astype(int))
    say("srgb exact fraction", (diff == 0).mean())
    assert diff.max() <= 1 and (diff == 0).mean() > 0.999
    rtx.pp_none(src.ptr, dst.ptr, w, h)
    rtx.synchronize()
    assert np.array_equal(dst.download(np.uint8, (h, w, 4)), orc.pp_none(raw))
    src.free(); dst.free()


# ---- the C++ batch driver over the same library ------------------------------------
@pytest.mark.parametrize("two_level,recipe", [(False, 0), (True, 0), (False, "glibc")])
def test_cpp_rtwo_driver_matches_python_path(rtx, two_level, recipe):
    """rtxplay_amd/rtwo (include/rtx/rtx.hpp over the C-ABI) renders the same
    bytes as the Python mirror for the same scene: seeded, flat and (-I)
    two-level, and on the reference's own rand() stream (-r glibc)."""
    import subprocess
    from conftest import ROOT
    from rtxplay_amd.scene import RTOWScene
    exe = os.path.join(ROOT, "rtxplay_amd", "rtwo")
    w, h, spp, depth = 48, 27, 3, 16
    say("rtwo driver", w, h, spp, "two-level", two_level)
    p = subprocess.run([exe, "-g", "%dx%d" % (w, h), "-s", str(spp), "-d", str(depth), "-l", "2", "-A", "RPP", "-S",
                        "-r", str(recipe)] + (["-I"] if two_level else []), capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    tok = p.stdout.split()
    assert tok[0] == "P3" and int(tok[1]) == w and int(tok[2]) == h
    ppm = np.array(tok[4:4 + 3 * w * h], np.int64).reshape(h, w, 3)[::-1]
    k = 4 + 3 * w * h
    assert tok[k] == "P2"
    pgm = np.array(tok[k + 4:k + 4 + w * h], np.int64).reshape(h, w)[::-1]
    assert "(pixel, rays, milliseconds)" in p.stderr
    s = RTOWScene(seed=recipe, half=11, level=2, two_level=two_level)
    assert s.load() == (486 if recipe == "glibc" else len(s.things_spec))
    s.build()
    L, lp = _launch(rtx, s, w, h, spp, depth)
    out = rtx.DeviceBuffer(4 * w * h)
    rtx.pp_sRGB(lp.rawRGB, out.ptr, w, h)
    rtx.synchronize()
    img = out.download(np.uint8, (h, w, 4))
    assert np.array_equal(ppm, img[..., :3].astype(np.int64))
    assert np.array_equal(pgm, L.rays_per_pixel().astype(np.int64))
    out.free(); L.free(); s.close()


def test_cpp_rtwo_reference_options(rtx):
    """The reference's batch options (optx/args.cxx:23-121): a named
    resolution with -a, long options, -A with a list, and the guides that
    -G prints only together with -D (optx/rtwo.cxx:549-558)."""
    import subprocess
    from conftest import ROOT
    from rtxplay_amd.scene import RTOWScene
    exe = os.path.join(ROOT, "rtxplay_amd", "rtwo")
    say("rtwo -g CGA -a 16:9")
    p = subprocess.run([exe, "-g", "CGA", "-a", "16:9", "-s", "1", "-d", "4", "-l", "2", "-q", "-S"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout == ""
    assert p.stderr.split()[0] == str(320 * 180)              # 320 x int(320 * 9 / 16 + .5)
    w, h, spp = 40, 24, 2
    say("rtwo long options, -D NRM -G -A RPP,XYZ")
    p = subprocess.run([exe, "--geometry", "%dx%d" % (w, h), "--samples-per-pixel", str(spp), "--trace-depth", "16",
                        "-l", "2", "--apply-denoiser", "NRM", "--print-guides", "--print-aov", "RPP,XYZ"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "unknown argument for option A ignored -- XYZ" in p.stderr
    tok = p.stdout.split()
    imgs, k = [], 0
    while k < len(tok):
        magic, iw, ih, mx = tok[k], int(tok[k + 1]), int(tok[k + 2]), int(tok[k + 3])
        n = iw * ih * (3 if magic == "P3" else 1)
        imgs.append((magic, mx, np.array(tok[k + 4:k + 4 + n], np.int64)))
        k += 4 + n
    assert [m for m, _, _ in imgs] == ["P3", "P3", "P3", "P2"]
    s = RTOWScene(seed=0, half=11, level=2)
    s.load()
    s.build()
    L, lp = _launch(rtx, s, w, h, spp, 16)
    want_n = (np.clip(L.normals(), 0, 1) * 255).astype(np.int64)[::-1].reshape(-1)
    want_a = (np.clip(L.albedos(), 0, 1) * 255).astype(np.int64)[::-1].reshape(-1)
    assert np.array_equal(imgs[0][2], want_n) and np.array_equal(imgs[1][2], want_a)
    assert np.array_equal(imgs[3][2], L.rays_per_pixel()[::-1].reshape(-1).astype(np.int64))
    # without -D the guides are not printed
    p = subprocess.run([exe, "-g", "%dx%d" % (w, h), "-s", "1", "-l", "2", "-G"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0 and p.stdout.split().count("P3") == 1
    L.free(); s.close()
