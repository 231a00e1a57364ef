"""Closest hits of small scenes, where the builder's 4-wide grouping works at
its edges: 2 to 40 triangles (the root stays a 4-wide node even where one
leaf would be cheaper; subtrees of a few triangles merge into one leaf) and
scenes of coincident copies (every hit a tie, resolved by the lower
primitive id).  Flat and two-level layouts.

* primitive ids against the FP64 brute force of the oracle wherever
  decidable, and t within a few ulps once scaled by the angle to the normal;
* against the oracle's FP32 traversal, which has the same watertight
  arithmetic and tie rule: the same primitive and t for (nearly) every ray.
"""
import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _soup(n, seed, copies=1):
    g = np.random.default_rng(seed)
    c = g.uniform(-1, 1, (n, 1, 3))
    t = (c + g.uniform(-.4, .4, (n, 3, 3))).astype(np.float32)
    return np.repeat(t, copies, axis=0)


def _rays_at(tris, n, seed):
    """Rays from a shell around the soup: 80 % aimed at random interior
    points of random triangles, the rest at random points of its box."""
    g = np.random.default_rng(seed)
    o = g.normal(size=(n, 3))
    o = 4.0 * o / np.linalg.norm(o, axis=1)[:, None]
    k = g.integers(0, len(tris), n)
    b = g.dirichlet((1.0, 1.0, 1.0), n)
    p = (b[:, :, None] * tris[k].astype(np.float64)).sum(axis=1)
    box = g.uniform(-1.4, 1.4, (n, 3))
    d = np.where((g.random(n) < 0.8)[:, None], p, box) - o
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-3
    rays[:, 4:7] = d
    rays[:, 7] = 1e16
    return rays


def _build(rtx, tris, two_level):
    s = rtx.Scene()
    if two_level:
        s.build_mode(rtx.Scene.TWO_LEVEL)
    verts = tris.reshape(-1, 3)
    idx = np.arange(len(verts), dtype=np.uint32).reshape(-1, 3)
    m = s.add((verts, idx))
    s.add(rtx.Thing.diffuse((.5, .5, .5)), m)
    s.build()
    return s


@pytest.mark.parametrize("two_level", [False, True])
@pytest.mark.parametrize("n,copies", [(2, 1), (3, 1), (5, 1), (8, 1), (9, 1), (17, 1), (40, 1), (4, 2), (3, 6)])
def test_small_scene_hits(rtx, orc, n, copies, two_level):
    tris = _soup(n, 1000 + 31 * n + copies, copies)
    say("scene of %d triangles (%d distinct), %s" % (len(tris), n, "two-level" if two_level else "flat"))
    s = _build(rtx, tris, two_level)
    try:
        world, inst = s.triangles()
        assert world.shape[0] == len(tris)
        rays = _rays_at(tris, 8000, 7 + n)
        hits = s.trace(rays)
        assert np.isfinite(hits["t"]).mean() > 0.5
        gp = np.where(hits["instance_id"] == 0xFFFFFFFF, 0xFFFFFFFF, hits["prim"]).astype(np.uint32)
        if copies == 1:
            ref64 = orc.closest_f64_brute(world, rays)
            ok = orc.decidable(ref64, rays, world)
            say("decidable fraction %.5f" % ok.mean())
            assert ok.mean() > 0.8
            assert not (ok & (gp != ref64["prim"])).any()
            # random orientations give grazing hits, where t is ill-conditioned:
            # its error times |cos| of the angle to the normal stays at a few ulps
            hit = ok & np.isfinite(ref64["t"])
            T = world.reshape(-1, 3, 3).astype(np.float64)[ref64["prim"][hit]]
            nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
            nrm /= np.linalg.norm(nrm, axis=1)[:, None]
            d = rays[hit, 4:7].astype(np.float64)
            d /= np.linalg.norm(d, axis=1)[:, None]
            cos = np.abs((nrm * d).sum(axis=1))
            rel = np.abs(hits["t"][hit].astype(np.float64) - ref64["t"][hit]) / ref64["t"][hit]
            assert (rel * cos).max() <= 1e-6
        ref = orc.BVH(world).closest_f32(rays)
        same = gp == ref["prim"]
        say("fp32 agreement", same.mean())
        assert same.mean() > 0.9995
        h = same & (gp != 0xFFFFFFFF)
        if not two_level:
            # the flat layout intersects in world space, as the oracle does
            assert np.array_equal(hits["t"][h], ref["t"][h])
        if copies > 1:
            # every hit is a tie among the copies: the lowest id of its group
            assert (gp[gp != 0xFFFFFFFF] % copies == 0).all()
    finally:
        s.close()
