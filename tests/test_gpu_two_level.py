"""Two-level scenes (RTX_BUILD_TWO_LEVEL): one BVH per mesh over its
object-space triangles under one BVH over the instances, rays transformed
into object space on entry, as the reference's IAS over GASes
(optx/scene.cxx:38-294); Scene::update is a refit of the top level.

What pins it:
* closest hits against the FP64 brute force over the world-space triangles
  (the flattening arithmetic, itself bit-exact against the oracle) on every
  decidable ray, with rotated, non-uniformly scaled, singular and empty
  instances;
* the refit: the closest hit is the minimum over the same object-space tests
  with an order-free tie rule, so after update() every hit is bit-identical
  to a fresh build of the moved scene;
* renders against the flat build of the same scene: the same Frand48
  streams, with hit points that differ in the last bits (object-space
  barycentrics), so images agree to MC noise, and all but a few pixels
  agree to 1e-3 -- a wrong transform, instance or normal breaks both.  The
  few: a scattered ray starts at the hit point rebuilt from the world
  vertices (optx/optics_r.cu:59-90), which lies off the surface the next
  trace tests by the rounding of world coordinates; leaving at a grazing
  angle it can hit that surface again beyond tmin = 1e-3.  Both layouts
  do this, at different rays, and only the ground sphere makes it common:
  it is the unit sphere scaled by 1000 (as the reference sets it up,
  optx/rtwo.cxx:151-163), so near its top the object-space ray origin
  rounds at 2^-23 x 1000 = 1.2e-4 world units, where world coordinates
  near y = 0 round far finer.  96 % of the pixels of the full level-4 scene
  agree to 1e-3 and the two-level frame traces about 1 % more rays (the
  extra bounces), 99.9 % agree without the ground;
* batches of two-level frames bit-identical to their plain launches, the
  picker, and the adjoint against finite differences of the render."""
import numpy as np
import pytest

from test_gpu_parity import _oracle_scene, _rays

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-2l]", *a, flush=True)


def _build(level, two_level, seed=0, half=11):
    from rtxplay_amd.scene import RTOWScene
    s = RTOWScene(seed=seed, half=half, level=level, two_level=two_level)
    s.load()
    say("build level", level, "half", half, "two-level", two_level)
    s.build()
    return s


def _normals(tris, prim):
    T = tris[prim].reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    return n / np.linalg.norm(n, axis=1)[:, None]


def _check_hits(s, tris, inst, rays, ref64, want_ratio=0.95, origin_inst=None):
    """Closest hits against the FP64 brute force over the world triangles.

    Primitive (and so instance) ids must agree on every decidable ray.  t:
    a two-level scene intersects exact transforms of the object-space
    triangles, the world triangles are those transforms rounded to float, so
    they differ by the rounding of world coordinates of the instance's
    magnitude (2^-24 of 1000-2000 for the ground sphere, 6e-5 world units),
    which moves a hit along the ray by that over the sine of the grazing
    angle.  t agrees to 1e-5 relative, or to 2^-15 of the ray's coordinate
    magnitude, or the hit moves across the triangle's plane by at most 2^-18
    of the largest world coordinate of the hit instance.

    Rays that start on a surface (origin_inst: the instance they start on)
    may hit that instance again within that band, where the world triangle
    and the transformed object-space one lie on either side of the origin;
    such self-hits are excused and counted."""
    hits = s.trace(rays)
    ok = __import__("oracle").decidable(ref64, rays, tris)
    say("decidable fraction %.5f of %d rays" % (ok.mean(), len(rays)))
    assert ok.mean() > want_ratio
    inst_mag = np.zeros(int(inst.max()) + 1 if len(inst) else 1)
    np.maximum.at(inst_mag, inst, np.abs(tris.reshape(len(tris), 9)).max(axis=1))
    h = np.isfinite(hits["t"])
    assert np.array_equal(hits["instance_id"][h], inst[hits["prim"][h]])
    if origin_inst is not None:
        self_hit = np.zeros(len(rays), bool)
        sh = h & (hits["instance_id"] == origin_inst)
        across = hits["t"][sh] * np.abs((rays[sh, 4:7].astype(np.float64) * _normals(tris, hits["prim"][sh])).sum(axis=1))
        self_hit[np.where(sh)[0]] = across <= 2.0 ** -18 * inst_mag[origin_inst[sh]]
        say("self-hits within the rounding band: %d of %d" % ((self_hit & ok & (hits["prim"] != ref64["prim"])).sum(),
                                                              len(rays)))
        ok = ok & ~self_hit
    gp = np.where(hits["instance_id"] == 0xFFFFFFFF, 0xFFFFFFFF, hits["prim"]).astype(np.uint32)
    bad = ok & (gp != ref64["prim"])
    assert not bad.any(), (np.where(bad)[0][:10], gp[bad][:10], ref64["prim"][bad][:10])
    m = ok & np.isfinite(ref64["t"])
    t32 = hits["t"][m].astype(np.float64)
    t64 = ref64["t"][m]
    r = rays[m].astype(np.float64)
    dn = np.linalg.norm(r[:, 4:7], axis=1)
    p = r[:, 0:3] + t64[:, None] * r[:, 4:7]
    mag = np.maximum(np.maximum(np.abs(r[:, 0:3]).max(axis=1), np.abs(p).max(axis=1)), 1.0)
    err = np.abs(t32 - t64)
    im = inst_mag[inst[ref64["prim"][m]]]
    across = err * np.abs((r[:, 4:7] * _normals(tris, ref64["prim"][m])).sum(axis=1))
    good = (err <= 1e-5 * t64) | (err * dn <= 2.0 ** -15 * mag) | (across <= 2.0 ** -18 * im)
    say("t within 1e-5 relative: %.5f, within the bound: %.5f" % ((err <= 1e-5 * t64).mean(), good.mean()))
    assert good.all(), (err[~good][:5], t64[~good][:5], im[~good][:5])
    return hits, ok


def _same_hits(a, b):
    for k in ("t", "u", "v", "instance_id", "primitive_index", "prim"):
        if not np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)):
            return False
    return True


@pytest.fixture(scope="module")
def small_tl(rtx, orc):
    s = _build(2, True, half=3)
    tris, inst = _oracle_scene(orc, s)
    yield s, tris, inst
    s.close()


@pytest.fixture(scope="module")
def level4_pair(rtx, orc):
    f = _build(4, False)
    t = _build(4, True)
    tris, inst = _oracle_scene(orc, f)
    yield f, t, tris, inst
    f.close()
    t.close()


def test_two_level_stats(small_tl):
    s, tris, _ = small_tl
    st = s.stats()
    say(st)
    assert st["two_level"] == 1 and st["num_blas"] == 1
    assert st["unique_prims"] == 64 and st["num_prims"] == len(tris) == 64 * len(s)
    assert st["tlas_nodes"] >= 1 and st["wide_nodes"] > st["tlas_nodes"]
    assert s.build_mode() == s.TWO_LEVEL


def test_two_level_triangles_are_the_flattened_ones(small_tl):
    s, tris, inst = small_tl
    t, i = s.triangles()
    assert np.array_equal(t.reshape(len(t), 9).view(np.uint32), tris.reshape(len(tris), 9).view(np.uint32))
    assert np.array_equal(i, inst)


def test_two_level_trace_vs_fp64_brute_force(small_tl, orc):
    s, tris, inst = small_tl
    rays = _rays(20000, 5)
    ref = orc.closest_f64_brute(tris, rays)
    _check_hits(s, tris, inst, rays, ref)


def test_two_level_camera_and_bounce_rays(level4_pair, orc):
    f, t, tris, inst = level4_pair
    cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 16 / 9, 0.1, 10.0)
    prim = orc.primary_rays(cam, 320, 180)
    bvh = orc.BVH(tris)
    ref = bvh.closest_f64(prim)
    _check_hits(t, tris, inst, prim, ref)
    # the same rays through the flat build: same primitive on every decidable ray
    hf, ht = f.trace(prim), t.trace(prim)
    ok = orc.decidable(ref, prim, tris)
    assert np.array_equal(hf["prim"][ok], ht["prim"][ok])
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
    _check_hits(t, tris, inst, b, bvh.closest_f64(b), want_ratio=0.9, origin_inst=inst[ref["prim"][h]])


def _xform(rot_axis, angle, scale, trans):
    a = np.asarray(rot_axis, np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K
    M = R @ np.diag(scale)
    return np.hstack([M, np.asarray(trans, np.float64)[:, None]]).astype(np.float32)


def _odd_scene(rtx, two_level):
    """Rotated, non-uniformly scaled and translated instances of a sphere and
    of a lopsided mesh, a singular instance (scale 0 on one axis) and an
    instance of an empty mesh."""
    s = rtx.Scene()
    if two_level:
        s.build_mode(s.TWO_LEVEL)
    v, i = rtx.Sphere(1.0, 3).mesh()
    ms = s.add((v, i))
    g = np.random.default_rng(3)
    lv = g.normal(size=(40, 3)).astype(np.float32) * np.float32([1.5, .4, .8])
    li = g.integers(0, 40, (60, 3)).astype(np.uint32)
    li = li[(li[:, 0] != li[:, 1]) & (li[:, 1] != li[:, 2]) & (li[:, 0] != li[:, 2])]
    ml = s.add((lv, li))
    me = s.add((np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)))
    xs = []
    for k in range(24):
        m = (ms, ml)[k % 2]
        t = s.add(rtx.Thing.diffuse((.5, .5, .5)), m)
        xf = _xform(g.normal(size=3), g.uniform(0, 2 * np.pi), g.uniform(.3, 1.5, 3), g.uniform(-5, 5, 3))
        s.set(t, xf)
        xs.append(xf)
    t = s.add(rtx.Thing.diffuse((.5, .5, .5)), ms)
    s.set(t, _xform((0, 1, 0), .3, (1., 0., 1.), (0., 0., 0.)))   # singular
    t = s.add(rtx.Thing.diffuse((.5, .5, .5)), me)                  # no triangles
    s.build()
    return s


def test_two_level_transforms_singular_and_empty(rtx, orc):
    s = _odd_scene(rtx, True)
    f = _odd_scene(rtx, False)
    tris, inst = f.triangles()
    t2, i2 = s.triangles()
    assert np.array_equal(t2.view(np.uint32), tris.view(np.uint32)) and np.array_equal(i2, inst)
    rays = _rays(20000, 11, lo=(-7, -7, -7), hi=(7, 7, 7))
    ref = orc.closest_f64_brute(tris, rays)
    hits, ok = _check_hits(s, tris, inst, rays, ref, want_ratio=0.9)
    singular = len(s) - 2
    assert not (hits["instance_id"] == singular).any()
    hf = f.trace(rays)
    assert np.array_equal(hf["prim"][ok], hits["prim"][ok])
    s.close()
    f.close()


def test_two_level_refit_equals_fresh_build(rtx, orc):
    from rtxplay_amd.scene import RTOWScene, transform_for
    s = _build(3, True, half=4)
    h = s.handle
    g = np.random.default_rng(5)
    moved = [dict(t) for t in s.things_spec]
    for k in range(1, len(moved), 3):
        c = np.array(moved[k]["center"], np.float32)
        moved[k]["center"] = tuple(np.float32(x) for x in c + g.uniform(-.6, .6, 3).astype(np.float32))
        s.set(k, transform_for(moved[k]))
    say("refit after moving", len(range(1, len(moved), 3)), "instances")
    assert s.update(h) == h
    st = s.stats()
    say("refit ms %.3f" % st["build_ms"])
    fresh = RTOWScene(seed=0, half=4, level=3, spec=moved, two_level=True)
    fresh.load()
    fresh.build()
    rays = _rays(30000, 17)
    cam = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 16 / 9, 0.1, 10.0)
    rays = np.concatenate([rays, orc.primary_rays(cam, 160, 90)])
    a, b = s.trace(rays), fresh.trace(rays)
    assert _same_hits(a, b)
    # and the hits are those of the moved triangles
    tris, inst = fresh.triangles()
    ref = orc.closest_f64_brute(tris, rays[:5000])
    _check_hits(s, tris, inst, rays[:5000], ref)
    # a thing added after the build: update rebuilds and the handle stays
    v, i = rtx.Sphere(1.0, 2).mesh()
    m = s.add((v, i))
    t = s.add(rtx.Thing.diffuse((.2, .2, .2)), m)
    s.set(t, np.float32([[1, 0, 0, 0], [0, 1, 0, 30], [0, 0, 1, 0]]))
    assert s.update(h) == h
    up = np.float32([[0, 25, 0, 1e-3, 0, 1, 0, 1e16]])
    hit = s.trace(up)[0]
    assert hit["instance_id"] == t and abs(hit["t"] - 4) < 1e-3
    fresh.close()
    s.close()


def _render(rtx, s, w, h, spp, depth):
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, s.handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    L = rtx.Launcher(s, lp)
    say("render %dx%d spp %d" % (w, h, spp))
    L.ignite()
    out = (L.raw_rgb(), L.rays_per_pixel(), L.normals(), L.albedos())
    L.free()
    return out


def _compare_renders(a, b, spp, what, close_frac=0.97, rpp_tol=0.01):
    raw_a, rpp_a, n_a, al_a = a
    raw_b, rpp_b, n_b, al_b = b
    w = raw_a.shape[1] * raw_a.shape[0]
    # per-channel means within the MC error (the per-pixel spread of the
    # difference bounds the error of the mean)
    d = (raw_a - raw_b).reshape(-1, 3).astype(np.float64)
    sigma = np.sqrt(((raw_a.reshape(-1, 3).astype(np.float64) * (1 - raw_a.reshape(-1, 3))) / spp).sum(axis=0)) / w
    dm = np.abs(d.mean(axis=0))
    close = np.abs(raw_a - raw_b).max(axis=2) < 1e-3
    say(what, "channel mean diff", dm, "3 sigma", 3 * sigma, "pixels within 1e-3", close.mean(),
        "rpp equal", (rpp_a == rpp_b).mean())
    assert (dm <= 3 * np.sqrt(2) * sigma + 1e-6).all()
    assert close.mean() >= close_frac
    assert (np.abs(al_a - al_b).max(axis=2) < 1e-3).mean() >= close_frac
    assert (np.abs(n_a - n_b).max(axis=2) < 1e-3).mean() >= close_frac
    assert abs(int(rpp_a.sum()) - int(rpp_b.sum())) / rpp_b.sum() < rpp_tol


def test_two_level_render_matches_flat(rtx, level4_pair):
    f, t, _, _ = level4_pair
    a = _render(rtx, t, 96, 54, 16, 16)
    b = _render(rtx, f, 96, 54, 16, 16)
    _compare_renders(a, b, 16, "level 4", close_frac=0.9, rpp_tol=0.025)
    raw = a[0]
    assert raw.min() >= 0 and raw.max() <= 1


def test_two_level_tuning_does_not_change_pixels(rtx, level4_pair):
    """Instance transitions run when enough of a wave waits for one, and the
    world ray and marker go to the LDS part of the stack or its global
    overflow area, so the launch tuning changes when and where transitions
    happen.  The closest hit does not depend on it: every tuning gives the
    same frame bit for bit (a 16-entry LDS stack sends the transitions of
    the deeper instances through the overflow area)."""
    _, t, _, _ = level4_pair
    old = t.tuning()
    ref = _render(rtx, t, 80, 45, 8, 16)
    try:
        for tun in (dict(stack_lds=16), dict(stack_lds=16, shade_batch=8, inner_frac=48),
                    dict(shade_batch=64, inner_frac=1), dict(blocks_per_cu=1, tile_w=64, shade_batch=1)):
            t.tuning(**{k: old[k] for k in ("tile_w", "stack_lds", "blocks_per_cu", "shade_batch", "inner_frac")})
            t.tuning(**tun)
            say("tuning", t.tuning())
            got = _render(rtx, t, 80, 45, 8, 16)
            for a, b in zip(got, ref):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), tun
    finally:
        t.tuning(**{k: old[k] for k in ("tile_w", "stack_lds", "blocks_per_cu", "shade_batch", "inner_frac")})


def test_two_level_render_matches_flat_without_ground(rtx, level4_pair):
    """Without the radius-1000 ground sphere the coordinates round at 1e-6
    and almost every path is the same in both layouts."""
    from rtxplay_amd.scene import RTOWScene
    f, _, _, _ = level4_pair
    spec = f.things_spec[1:]
    out = []
    for tl in (True, False):
        s = RTOWScene(level=4, spec=spec, two_level=tl)
        s.load()
        s.build()
        out.append(_render(rtx, s, 96, 54, 16, 16))
        s.close()
    _compare_renders(out[0], out[1], 16, "level 4 without ground", close_frac=0.995, rpp_tol=0.002)


def test_two_level_picker_matches_flat(rtx, level4_pair):
    f, t, _, _ = level4_pair
    w, h = 320, 180
    ids = []
    for s in (f, t):
        lp = rtx.LpGeneral()
        lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, 1, 16, s.handle
        lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
        L = rtx.Launcher(s, lp)
        got = []
        for (x, y) in [(160, 90), (10, 170), (300, 20), (5, 5), (200, 120)]:
            lp.picker, lp.pick_x, lp.pick_y = True, x, y
            L.ignite(once=True)
            got.append(L.pick_id())
        L.free()
        ids.append(got)
    say("picks", ids)
    assert ids[0] == ids[1]


def test_two_level_mixed_reference_scene(rtx):
    """The reference's own scene: per-sphere levels 9/8/6/3 (4 meshes,
    8.79 M instanced triangles from 1.33 M stored ones)."""
    t = _build("mixed", True)
    st = t.stats()
    say(st)
    assert st["num_blas"] == 4
    assert st["unique_prims"] == 4 ** 10 + 4 ** 9 + 4 ** 7 + 4 ** 4
    levels = t.levels
    assert st["num_prims"] == sum(4 * 4 ** lv for lv in levels)
    a = _render(rtx, t, 128, 72, 8, 16)
    t.close()
    f = _build("mixed", False)
    b = _render(rtx, f, 128, 72, 8, 16)
    f.close()
    _compare_renders(a, b, 8, "mixed", close_frac=0.9, rpp_tol=0.025)


def test_two_level_batch_equals_plain_launches(rtx):
    from rtxplay_amd.scene import BatchFrames
    from rtxplay_amd.hip import Stream
    from test_gpu_batch import _frames, _outputs, _clear, _same
    bf = BatchFrames(_frames(), two_level=True)
    for l in bf.launchers:
        l.ignite()
    ref = _outputs(bf)
    for use_graph in (False, True):
        _clear(bf)
        b = rtx.Batch(bf.scenes[0], bf.lps, use_graph=use_graph)
        st = Stream()
        say("two-level batch launch, graph", use_graph)
        b.launch(st)
        b.launch(st)
        st.synchronize()
        for i, (g, r) in enumerate(zip(_outputs(bf), ref)):
            assert _same(g, r), "frame %d differs (graph=%s)" % (i, use_graph)
        b.free()
    # flat and two-level scenes do not mix in one batch
    from rtxplay_amd.scene import RTOWScene
    flat = RTOWScene(seed=9, half=3, level=2)
    flat.load()
    lp = bf.lps[0]
    keep = lp.is_handle
    lp.is_handle = flat.build()
    with pytest.raises(rtx.RtxError):
        rtx.Batch(bf.scenes[0], [bf.lps[1], lp], use_graph=False)
    lp.is_handle = keep
    flat.close()
    bf.close()


def test_batch_follows_a_rebuilt_scene(rtx):
    """rtx_scene_update of a flat scene replaces its arrays; a batch created
    before picks the new scene up at its next launch."""
    from rtxplay_amd.scene import BatchFrames, transform_for
    from rtxplay_amd.hip import Stream
    from test_gpu_batch import _frames, _outputs, _same
    bf = BatchFrames(_frames()[:3])
    b = rtx.Batch(bf.scenes[0], bf.lps, use_graph=True)
    st = Stream()
    b.launch(st)
    st.synchronize()
    sc = bf.scenes[1]
    spec = sc.things_spec
    for k in range(1, len(spec), 2):
        t = dict(spec[k])
        t["center"] = (t["center"][0], np.float32(t["center"][1] + .3), t["center"][2])
        sc.set(k, transform_for(t))
    say("update a flat scene under a captured batch")
    sc.update(sc.handle)
    b.launch(st)
    st.synchronize()
    got = _outputs(bf)
    for l in bf.launchers:
        l.ignite()
    assert _same(got[1], _outputs(bf)[1])
    b.free()
    bf.close()


def test_two_level_adjoint_matches_finite_differences(rtx):
    """Albedo never changes a path, so central differences of the GPU render
    pin the adjoint of a two-level scene as they do a flat one's."""
    s = _build(3, True, half=5)
    w, h, spp, depth = 96, 54, 16, 8
    lp = rtx.LpGeneral()
    lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = w, h, spp, depth, s.handle
    lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, w / h, 0.1, 10.0).c
    L = rtx.Launcher(s, lp)
    m0 = s.get_materials()
    T = np.clip(m0 * np.float32(.8), 0, 1)
    s.set_materials(T)
    L.ignite()
    target = L.raw_rgb().astype(np.float64)
    s.set_materials(m0)
    L.ignite()
    img = L.raw_rgb()
    g = L.adjoint((2.0 * (img.astype(np.float64) - target)).astype(np.float32)).astype(np.float64)

    def loss(m):
        s.set_materials(m)
        L.ignite()
        return float(((L.raw_rgb().astype(np.float64) - target) ** 2).sum())

    cand = sorted(((abs(g[k, c]), k, c) for k in range(len(g)) for c in range(3)), reverse=True)[:6]
    for _, k, c in cand:
        mp, mm = m0.copy(), m0.copy()
        mp[k, c] = np.float32(m0[k, c] + 1e-3)
        mm[k, c] = np.float32(m0[k, c] - 1e-3)
        fd = (loss(mp) - loss(mm)) / (float(mp[k, c]) - float(mm[k, c]))
        rel = abs(g[k, c] - fd) / abs(fd)
        say("albedo", k, c, "adjoint %.6e fd %.6e rel %.2e" % (g[k, c], fd, rel))
        assert rel <= 1e-3
    s.set_materials(m0)
    L.free()
    s.close()


_DET_SCRIPT = r'''
import hashlib, sys
sys.path.insert(0, sys.argv[1])
import rtxplay_amd as rtx
from rtxplay_amd.scene import RTOWScene
s = RTOWScene(seed=0, half=11, level=4, two_level=True)
s.load()
s.build()
lp = rtx.LpGeneral()
lp.image_w, lp.image_h, lp.spp, lp.depth, lp.is_handle = 96, 54, 16, 16, s.handle
lp.camera = rtx.Camera().set((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 96 / 54, 0.1, 10.0).c
L = rtx.Launcher(s, lp)
L.ignite()
print(hashlib.sha1(L.raw_rgb().tobytes() + L.rays_per_pixel().tobytes()).hexdigest())
'''


def test_two_level_frame_is_the_same_in_every_process(rtx):
    """The frame does not depend on which lanes take which pixels (the work
    counter hands them out in a different order every launch): two fresh
    processes render the same bytes."""
    import subprocess
    import sys
    from conftest import ROOT
    outs = []
    for k in range(2):
        say("fresh process", k)
        p = subprocess.run([sys.executable, "-c", _DET_SCRIPT, ROOT], capture_output=True, text=True, timeout=180)
        assert p.returncode == 0, p.stderr
        outs.append(p.stdout.strip().split()[-1])
    assert outs[0] == outs[1]
