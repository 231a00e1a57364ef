"""Pins of the oracle (CPU restatement) against the reference's own outputs,
glibc, closed forms and invariants.  No GPU."""
import hashlib
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


# ---- the reference CPU tracer itself -----------------------------------------
@pytest.mark.slow
def test_rtow_restatement_bit_exact(orc):
    """orc_rtow_render reproduces rtow.cxx's PPM byte for byte (reference binary
    built from rtow.cxx, golden written by tests/golden/make_rtow_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "rtow_ref.npz"))
    img, rays = orc.rtow_render(int(g["w"]), int(g["h"]), int(g["spp"]), int(g["depth"]), argorder=1)
    assert np.array_equal(img[g["rows"]], g["row_pixels"])
    assert np.array_equal(img.reshape(-1, 3).sum(axis=0, dtype=np.int64), g["channel_sums"])
    assert hashlib.sha256(img.tobytes()).digest() == bytes(g["sha256"])
    assert rays > img.shape[0] * img.shape[1] * int(g["spp"])


def test_rtow_ray_count_of_a_prefix(orc):
    """orc_rtow_rays (the benchmark's ray count for the reference tracer's
    finished rows) counts the rays of the first pixels of the same render:
    the whole image gives the render's count, a row more adds at least one
    ray per sample of that row, and the count does not depend on pixels
    after the prefix."""
    w, h, spp, depth = 48, 27, 2, 50
    _, rays = orc.rtow_render(w, h, spp, depth, argorder=1)
    assert orc.rtow_rays(w * h, w, h, spp, depth) == rays
    counts = [orc.rtow_rays(r * w, w, h, spp, depth) for r in range(h + 1)]
    assert counts[0] == 0 and counts[-1] == rays
    assert all(b - a >= w * spp for a, b in zip(counts, counts[1:]))
    assert orc.rtow_rays(5 * w, w, h, spp, depth) == orc.rtow_rays(5 * w, w, h, spp, depth)


def test_rtow_argorder_matters(orc):
    """The other argument-evaluation order gives another scene (so the exact
    match above really pins the order)."""
    a = orc.rtow_scene_glibc(1)
    b = orc.rtow_scene_glibc(0)
    assert len(a) > 400 and len(b) > 400
    ca = np.array([list(t.center) for t in a[:50]])
    cb = np.array([list(t.center) for t in b[:50]])
    assert not np.array_equal(ca, cb)


# ---- Frand48 -----------------------------------------------------------------
def test_frand48_glibc_kat(orc):
    k = np.load(os.path.join(GOLDEN, "frand48_kat.npz"))
    got = orc.frand48_draws(k["seeds"], k["draws"].shape[1])
    assert np.array_equal(got.view(np.uint32), k["draws"].view(np.uint32))


def test_frand48_closed_form(orc):
    """x_{n+1} = (a x_n + c) mod 2^48, out = RN_f32(x)/2^48, 16 discards."""
    seeds = np.array([0, 1, 77, 2 ** 31 + 5], np.uint32)
    got = orc.frand48_draws(seeds, 20)
    for r, s in enumerate(seeds):
        x = int(s)
        for _ in range(16):
            x = (0x5DEECE66D * x + 0xB) % (1 << 48)
        for j in range(20):
            x = (0x5DEECE66D * x + 0xB) % (1 << 48)
            assert got[r, j] == np.float32(x / 2.0 ** 48)


def test_frand48_can_return_one(orc):
    """Round-to-nearest of a 48-bit value can give exactly 1.0f (x >= 2^48 - 2^23)."""
    x = (1 << 48) - 1
    assert np.float32(float(x)) / np.float32(2.0 ** 48) == np.float32(1.0)


def test_host_frand48_matches_oracle(orc):
    from rtxplay_amd.scene import Frand48
    seeds = [0, 3, 4711, 2073599]
    want = orc.frand48_draws(np.array(seeds, np.uint32), 50)
    for r, s in enumerate(seeds):
        g = Frand48(s)
        got = np.array([g() for _ in range(50)], np.float32)
        assert np.array_equal(got.view(np.uint32), want[r].view(np.uint32))


# ---- camera -----------------------------------------------------------------
def _camera_f64(eye, pat, vup, fov, asp, aperture, foc):
    eye, pat, vup = (np.array(v, np.float64) for v in (eye, pat, vup))
    w = (eye - pat) / np.linalg.norm(eye - pat)
    u = np.cross(vup, w); u /= np.linalg.norm(u)
    v = np.cross(w, u)
    h = 2 * math.tan(0.5 * math.radians(fov))
    return dict(u=u, v=v, w=w, hvec=foc * h / 2 * v, wvec=foc * h * asp / 2 * u, dvec=foc * w)


@pytest.mark.parametrize("asp", [16 / 9, 1.0, 4 / 3])
def test_camera_closed_form(orc, asp):
    args = ((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, asp, 0.1, 10.0)
    c = orc.camera(*args)
    ref = _camera_f64(*args)
    for k, val in ref.items():
        got = np.array(list(getattr(c, k)), np.float64)
        assert np.allclose(got, val, rtol=1e-6, atol=1e-6), k


def test_camera_lens_sample_in_disk(orc):
    c = orc.camera((13, 2, 3), (0, 0, 0), (0, 1, 0), 20.0, 16 / 9, 0.1, 10.0)
    rays = orc.primary_rays(c, 64, 36)
    off = rays[:, 0:3] - np.array([13, 2, 3], np.float32)
    r = np.linalg.norm(off, axis=1)
    assert (r < 0.05 + 1e-6).all() and r.max() > 0.03
    # offsets lie in the lens plane spanned by u, v (perpendicular to w)
    w = np.array(list(c.w))
    assert np.abs(off @ w).max() < 1e-6


def test_primary_ray_pixel_mapping(orc):
    """s = 2(x+jitter)/w - 1 (optx/camera_r.cu:69-70): with aperture 0 the ray
    through pixel (x,y) points inside that pixel's footprint on the focus plane."""
    w, h = 32, 18
    c = orc.camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 90.0, w / h, 0.0, 1.0)
    rays = orc.primary_rays(c, w, h)
    d = rays[:, 4:7] / -rays[:, 6:7]   # intersect the plane z = -1
    ys, xs = np.divmod(np.arange(w * h), w)
    sx = (d[:, 0] / (w / h) + 1) / 2 * w
    sy = (d[:, 1] + 1) / 2 * h
    assert ((sx >= xs - 1e-4) & (sx <= xs + 1 + 1e-4)).all()
    assert ((sy >= ys - 1e-4) & (sy <= ys + 1 + 1e-4)).all()


# ---- sphere tessellation ----------------------------------------------------
def test_reduce_kat(orc):
    """Known answer of optx/reduce.cxx:66-105."""
    vt = [(0, 0, 0), (1, 0, 0), (.5, 1, 0), (1, 0, 0), (2, 0, 0), (1.5, 1, 0), (.5, 1, 0), (1.5, 1, 0), (1, 2, 0),
          (.5, 1, 0), (1.5, 1, 0), (1, 0, 0)]
    v, i = orc.reduce(vt)
    assert v.tolist() == [[0, 0, 0], [1, 0, 0], [.5, 1, 0], [2, 0, 0], [1.5, 1, 0], [1, 2, 0]]
    assert i.tolist() == [[0, 1, 2], [1, 3, 4], [2, 4, 5], [2, 4, 1]]


def test_reduce_signed_zero(orc):
    """-0 and +0 are one vertex (the reference's comparator uses ==)."""
    v, i = orc.reduce([(0.0, 1, 0), (-0.0, 1, 0), (1, 1, 1)])
    assert len(v) == 2 and i.tolist() == [[0, 0, 1]]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5])
def test_sphere_mesh_invariants(orc, n):
    v, i = orc.sphere_mesh(1.0, n)
    assert len(v) == 2 * 4 ** n + 2 and len(i) == 4 * 4 ** n
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
    edges = np.concatenate([i[:, [0, 1]], i[:, [1, 2]], i[:, [2, 0]]])
    # each directed edge once, each undirected edge in exactly two faces
    assert len({tuple(e) for e in edges.tolist()}) == len(edges)
    und = np.sort(edges, axis=1)
    _, counts = np.unique(und, axis=0, return_counts=True)
    assert (counts == 2).all()
    E = len(counts)
    assert len(v) - E + len(i) == 2           # Euler characteristic of a sphere
    a, b, c = v[i[:, 0]], v[i[:, 1]], v[i[:, 2]]
    nrm = np.cross(b - a, c - a)
    assert ((nrm * (a + b + c)).sum(axis=1) < 0).all()   # base winding is inward


def test_sphere_mesh_radius(orc):
    v, _ = orc.sphere_mesh(2.5, 3)
    assert np.allclose(np.linalg.norm(v, axis=1), 2.5, rtol=1e-6)


def test_vectorised_transform_matches_c(orc):
    """oracle.transform_points (numpy) == orc_transform_point (C), bit for bit."""
    g = np.random.default_rng(4)
    xf = g.normal(size=12).astype(np.float32)
    v = g.normal(size=(200, 3)).astype(np.float32)
    want = np.stack([orc.transform_point(xf, p) for p in v])
    assert np.array_equal(orc.transform_points(xf, v).view(np.uint32), want.view(np.uint32))


# ---- scene recipe -------------------------------------------------------------
def test_scene_recipe_matches_host(orc):
    from rtxplay_amd.scene import sphere_things
    for seed, half in [(0, 11), (7, 4), (63, 3)]:
        o = orc.scene_rtow(seed, half)
        p = sphere_things(seed, half)
        assert len(o) == len(p)
        for a, b in zip(o, p):
            assert a.type == b["type"] and a.kind == b["kind"]
            assert np.array_equal(np.array(list(a.albedo), np.float32), np.array(b["albedo"], np.float32))
            assert np.array_equal(np.array(list(a.center), np.float32), np.array(b["center"], np.float32))
            assert np.float32(a.fuzz) == b["fuzz"] and np.float32(a.index) == b["index"]
            assert np.float32(a.radius) == b["radius"]


def test_scene_recipe_shape(orc):
    s = orc.scene_rtow(0, 11)
    types = np.array([t.type for t in s])
    assert 450 < len(s) <= 488
    assert s[0].radius == 1000 and s[-1].type == orc.ORC_REFLECT and s[-3].type == orc.ORC_REFRACT
    small = types[1:-3]
    frac = np.bincount(small, minlength=3) / len(small)
    assert 0.7 < frac[0] < 0.9 and 0.08 < frac[1] < 0.22 and frac[2] < 0.1
    c = np.array([list(t.center) for t in s[1:-3]])
    assert (np.linalg.norm(c - np.array([4, .2, 0]), axis=1) > 0.9).all()


def test_scene_recipe_draw_ranges(orc):
    """The draws of optx/rtwo.cxx:165-207: a small sphere of radius .2 at
    (a + .9u, .2, b + .9u) for each grid cell a, b in [-11, 11) at most once;
    diffuse albedo = rnd3 * rnd3 in [0, 1), metal albedo in [.5, 1) and fuzz in
    [0, .5), glass index 1.5; the ground and the three big spheres as listed
    (optx/rtwo.cxx:154-163, 209-243)."""
    for seed in (0, 5, 17):
        s = orc.scene_rtow(seed, 11)
        g = s[0]
        assert (g.radius, tuple(g.center), g.type) == (1000, (0, -1000, 0), orc.ORC_DIFFUSE)
        assert np.array_equal(np.array(list(g.albedo), np.float32), np.float32([.5, .5, .5]))
        big = s[-3:]
        assert [(b.type, tuple(b.center), b.radius) for b in big] == [
            (orc.ORC_REFRACT, (0, 1, 0), 1), (orc.ORC_DIFFUSE, (-4, 1, 0), 1), (orc.ORC_REFLECT, (4, 1, 0), 1)]
        assert np.float32(big[0].index) == np.float32(1.5)
        assert np.array_equal(np.array(list(big[1].albedo), np.float32), np.float32([.4, .2, .1]))
        assert np.array_equal(np.array(list(big[2].albedo), np.float32), np.float32([.7, .6, .5]))
        assert big[2].fuzz == 0
        cells = set()
        for t in s[1:-3]:
            x, y, z = (float(v) for v in t.center)
            assert np.float32(t.radius) == np.float32(.2) and np.float32(y) == np.float32(.2)
            a, b = int(np.floor(x)), int(np.floor(z))
            assert -11 <= a < 11 and -11 <= b < 11 and (a, b) not in cells
            cells.add((a, b))
            assert 0 <= x - a <= .9 + 1e-5 and 0 <= z - b <= .9 + 1e-5
            alb = np.array(list(t.albedo), np.float64)
            if t.type == orc.ORC_DIFFUSE:
                assert (alb >= 0).all() and (alb < 1).all()
            elif t.type == orc.ORC_REFLECT:
                assert (alb >= .5).all() and (alb < 1).all() and 0 <= t.fuzz < .5
            else:
                assert t.type == orc.ORC_REFRACT and np.float32(t.index) == np.float32(1.5)
        # the rejection around (4, .2, 0) skips at most the few cells near it
        assert 480 <= len(cells) + 4 <= 488


# ---- closest hit --------------------------------------------------------------
def _ray(o, d, tmin=1e-3, tmax=1e16):
    return np.array(list(o) + [tmin] + list(d) + [tmax], np.float32)


def test_closest_f64_analytic(orc):
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    rays = np.stack([_ray((.25, .25, 1), (0, 0, -1)),       # hit at t=1, u=v=.25
                     _ray((.25, .25, 1), (0, 0, -2)),       # unnormalised: t=.5
                     _ray((2, 2, 1), (0, 0, -1)),           # miss outside
                     _ray((.25, .25, 1), (0, 0, 1)),        # pointing away
                     _ray((.25, .25, 1), (0, 0, -1), tmax=.5),  # beyond tmax
                     _ray((.25, .25, -1), (0, 0, 1))])      # from behind
    r = orc.closest_f64_brute(tri, rays)
    assert r["prim"].tolist() == [0, 0, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0]
    assert np.allclose(r["t"][[0, 1, 5]], [1.0, 0.5, 1.0])
    assert np.allclose([r["u"][0], r["v"][0]], [.25, .25])


def test_closest_f64_nearest_and_tie(orc):
    tris = np.array([[0, 0, -2, 1, 0, -2, 0, 1, -2],      # farther
                     [0, 0, -1, 1, 0, -1, 0, 1, -1],      # nearer
                     [0, 0, -1, 1, 0, -1, 0, 1, -1]], np.float32)   # exact duplicate of 1
    r = orc.closest_f64_brute(tris, _ray((.2, .2, 0), (0, 0, -1))[None])
    assert r["prim"][0] == 1 and r["t"][0] == 1.0 and r["t2"][0] == 1.0   # tie -> lower id, undecidable


def _small_scene(orc, level=2, seed=0, half=3):
    from rtxplay_amd.scene import sphere_things
    th = sphere_things(seed, half)
    meshes = {level: orc.sphere_mesh(1.0, level)}
    return orc.flatten(meshes, th, [level] * len(th)) + (th,)


def _seeded_rays(n, seed, lo=(-6, .1, -6), hi=(6, 3, 6)):
    g = np.random.default_rng(seed)
    o = g.uniform(lo, hi, (n, 3))
    d = g.normal(size=(n, 3))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o; rays[:, 3] = 1e-3; rays[:, 4:7] = d; rays[:, 7] = 1e16
    return rays


def test_bvh_f64_equals_brute_force(orc):
    tris, inst, th = _small_scene(orc)
    rays = _seeded_rays(3000, 1)
    a = orc.closest_f64_brute(tris, rays)
    b = orc.BVH(tris).closest_f64(rays)
    assert np.array_equal(a["prim"], b["prim"])
    assert np.array_equal(a["t"], b["t"]) and np.array_equal(a["t2"], b["t2"])


def test_watertight_f32_matches_f64_when_decidable(orc):
    tris, inst, th = _small_scene(orc)
    rays = _seeded_rays(20000, 2)
    a = orc.closest_f64_brute(tris, rays)
    b = orc.BVH(tris).closest_f32(rays)
    ok = orc.decidable(a, rays, tris)
    assert ok.mean() > 0.97
    assert np.array_equal(a["prim"][ok], b["prim"][ok])
    hit = ok & np.isfinite(a["t"])
    assert t_within_tolerance(b["t"], a["t"], rays, hit)


def t_within_tolerance(t32, t64, rays, mask):
    """FP32 hit distances: relative 1e-5, or within 2^-15 of the coordinate
    magnitude in world units (short hits far from the origin), and at least
    99.9% of hits within 1e-5 relative."""
    t32 = np.asarray(t32, np.float64)[mask]
    t64 = np.asarray(t64, np.float64)[mask]
    r = np.asarray(rays, np.float32)[mask].astype(np.float64)
    dn = np.linalg.norm(r[:, 4:7], axis=1)
    p = r[:, 0:3] + t64[:, None] * r[:, 4:7]
    mag = np.maximum(np.maximum(np.abs(r[:, 0:3]).max(axis=1), np.abs(p).max(axis=1)), 1.0)
    err = np.abs(t32 - t64)
    ok = (err <= 1e-5 * t64) | (err * dn <= 2.0 ** -15 * mag)
    return bool(ok.all()) and (err <= 1e-5 * t64).mean() >= 0.999


def test_watertight_no_leak_through_vertices_and_edges(orc):
    """Rays from the centre of a closed sphere mesh aimed exactly at every
    vertex and edge midpoint must hit (watertightness)."""
    v, i = orc.sphere_mesh(1.0, 3)
    tris = v[i].reshape(-1, 9)
    mids = np.concatenate([(v[i[:, 0]] + v[i[:, 1]]) / 2, (v[i[:, 1]] + v[i[:, 2]]) / 2])
    targets = np.concatenate([v, mids]).astype(np.float32)
    rays = np.zeros((len(targets), 8), np.float32)
    rays[:, 3] = 1e-3; rays[:, 4:7] = targets; rays[:, 7] = 1e16
    b = orc.BVH(tris).closest_f32(rays)
    assert (b["prim"] != 0xFFFFFFFF).all()


# ---- scatter building blocks (optx/optics_r.cu) -------------------------------
def test_scatter_kats(orc):
    s2 = np.float32(1 / math.sqrt(2))
    assert np.allclose(orc.reflect((1, -1, 0), (0, 1, 0)), (s2, s2, 0), atol=1e-7)
    assert np.allclose(orc.refract((1, -1, 0), (0, 1, 0), 1.0), (s2, -s2, 0), atol=1e-6)   # ratio 1: straight
    assert abs(orc.schlick(1.0, 1 / 1.5) - 0.04) < 1e-6
    assert abs(orc.schlick(0.0, 1 / 1.5) - 1.0) < 1e-6
    # Snell: sin(out) = ratio * sin(in)
    d = np.array([math.sin(0.5), -math.cos(0.5), 0], np.float32)
    out = orc.refract(d, (0, 1, 0), 1 / 1.5)
    assert abs(out[0] - math.sin(0.5) / 1.5) < 1e-6 and out[1] < 0
    assert np.allclose(orc.sky((0, 1, 0)), (.5, .7, 1.)) and np.allclose(orc.sky((0, -5, 0)), (1, 1, 1))
    assert np.allclose(orc.sky((1, 0, 0)), (.75, .85, 1.))


# ---- post-process -------------------------------------------------------------
def test_srgb_curve(orc):
    x = np.linspace(0, 1, 4097, dtype=np.float32).reshape(1, -1, 1).repeat(3, axis=2)
    u8 = orc.pp_srgb(x)
    ref = orc.pp_srgb_f64(x)
    assert (np.abs(u8[..., :3].astype(np.float64) - np.floor(ref)) <= 1).all()
    # FP32: 1.055f*1 - .055f = 0.99999994, so white truncates to 254 (as in the reference kernel)
    assert u8[0, 0, 0] == 0 and u8[0, -1, 0] == 254 and (u8[..., 3] == 255).all()
    assert orc.pp_none(x)[0, -1, 0] == 255 and orc.pp_none(x)[0, 2048, 0] == 127
